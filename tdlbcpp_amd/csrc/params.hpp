// JSON-serialisable parameter structs of the ComputeUnit: the reference's input
// and AllParams formats (tdlbcpp/src/Params/*.hpp).  Keys, value kinds and
// defaults follow those files; a missing or mistyped key throws
// std::runtime_error("Failed to parse <Struct>: ...") as the reference does.
#pragma once

#include <cstdint>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "json.hpp"

namespace tdlb {

using tNi = long int;         // Header.h:42
using tStep = unsigned long;  // Header.h:46

template <typename F>
void parse_guard(const char* name, F&& f) {
    try {
        f();
    } catch (const json::error& e) {
        throw std::runtime_error(std::string("Failed to parse ") + name + ": " + e.what());
    }
}

// GridParams.hpp:28-63
struct GridParams {
    tNi ngx = 1, ngy = 1, ngz = 1;
    tNi x = 60, y = 60, z = 60;
    tNi multiStep = 1;
    std::string strMinQVecPrecision = "float";

    void getParamsFromJson(const json::Value& j) {
        parse_guard("GridParams", [&] {
            ngx = (tNi)j.at("ngx").get<uint64_t>();
            ngy = (tNi)j.at("ngy").get<uint64_t>();
            ngz = (tNi)j.at("ngz").get<uint64_t>();
            x = (tNi)j.at("x").get<uint64_t>();
            y = (tNi)j.at("y").get<uint64_t>();
            z = (tNi)j.at("z").get<uint64_t>();
            multiStep = (tNi)j.at("multiStep").get<uint64_t>();
            strMinQVecPrecision = j.at("strMinQVecPrecision").get<std::string>();
        });
    }
    json::Value getJson() const {
        json::Value j = json::Value::object();
        j["ngx"] = (long long)ngx; j["ngy"] = (long long)ngy; j["ngz"] = (long long)ngz;
        j["x"] = (long long)x; j["y"] = (long long)y; j["z"] = (long long)z;
        j["multiStep"] = (long long)multiStep;
        j["strMinQVecPrecision"] = strMinQVecPrecision;
        return j;
    }
};

// FlowParams.hpp:27-105 (values held in double here; the ComputeUnit converts to T)
struct FlowParams {
    double initialRho = 8.0, reMNonDimensional = 7300.0, uav = 0.1, cs0 = 0.12, g3 = 0.8, nu = 0.0;
    double fx0 = 0.0, fy0 = 0.0, fz0 = 0.0, Re_m = 0.0, Re_f = 0.0, uf = 0.0, alpha = 0.97, beta = 1.9;
    bool useLES = false;
    std::string collision = "EgglesSomers";
    std::string streaming = "Esotwist";

    void getParamsFromJson(const json::Value& j) {
        parse_guard("FlowParams", [&] {
            initialRho = j.at("initialRho").get<double>();
            reMNonDimensional = j.at("reMNonDimensional").get<double>();
            uav = j.at("uav").get<double>();
            cs0 = j.at("cs0").get<double>();
            g3 = j.at("g3").get<double>();
            nu = j.at("nu").get<double>();
            fx0 = j.at("fx0").get<double>();
            fy0 = j.at("fy0").get<double>();
            fz0 = j.at("fz0").get<double>();
            Re_m = j.at("Re_m").get<double>();
            Re_f = j.at("Re_f").get<double>();
            uf = j.at("uf").get<double>();
            alpha = j.at("alpha").get<double>();
            beta = j.at("beta").get<double>();
            useLES = j.at("useLES").get<bool>();
            collision = j.at("collision").get<std::string>();
            streaming = j.at("streaming").get<std::string>();
        });
    }
    json::Value getJson() const {
        json::Value j = json::Value::object();
        j["initialRho"] = initialRho; j["reMNonDimensional"] = reMNonDimensional; j["uav"] = uav;
        j["cs0"] = cs0; j["g3"] = g3; j["nu"] = nu; j["fx0"] = fx0; j["fy0"] = fy0; j["fz0"] = fz0;
        j["Re_m"] = Re_m; j["Re_f"] = Re_f; j["uf"] = uf; j["alpha"] = alpha; j["beta"] = beta;
        j["useLES"] = useLES; j["collision"] = collision; j["streaming"] = streaming;
        return j;
    }
};

// RunningParams.hpp:26-68 (doubleResolutionAtStep defaults to 10 when absent)
struct RunningParams {
    tStep step = 1;
    double angle = 0.0;
    tStep num_steps = 20;
    tStep impellerStartupStepsUntilNormalSpeed = 30;
    tStep numStepsForAverageCalc = 10;
    tStep repeatPrintTimerToFile = 20;
    tStep repeatPrintTimerToStdOut = 10;
    std::string runningDataFileDir = ".";
    std::string runningDataFilePrefix = "debug";
    bool runningDataFileAppendTime = true;
    tStep doubleResolutionAtStep = 10;

    void incrementStep() { step++; }
    void getParamsFromJson(const json::Value& j) {
        parse_guard("RunningParams", [&] {
            step = j.at("step").get<tStep>();
            angle = j.at("angle").get<double>();
            num_steps = j.at("num_steps").get<tStep>();
            impellerStartupStepsUntilNormalSpeed = j.at("impellerStartupStepsUntilNormalSpeed").get<tStep>();
            numStepsForAverageCalc = j.at("numStepsForAverageCalc").get<tStep>();
            repeatPrintTimerToFile = j.at("repeatPrintTimerToFile").get<tStep>();
            repeatPrintTimerToStdOut = j.at("repeatPrintTimerToStdOut").get<tStep>();
            runningDataFileDir = j.at("runningDataFileDir").get<std::string>();
            runningDataFilePrefix = j.at("runningDataFilePrefix").get<std::string>();
            runningDataFileAppendTime = j.at("runningDataFileAppendTime").get<bool>();
            doubleResolutionAtStep = j.value<tStep>("doubleResolutionAtStep", 10);
        });
    }
    json::Value getJson() const {
        json::Value j = json::Value::object();
        j["step"] = (unsigned long long)step; j["angle"] = angle; j["num_steps"] = (unsigned long long)num_steps;
        j["impellerStartupStepsUntilNormalSpeed"] = (unsigned long long)impellerStartupStepsUntilNormalSpeed;
        j["numStepsForAverageCalc"] = (unsigned long long)numStepsForAverageCalc;
        j["repeatPrintTimerToFile"] = (unsigned long long)repeatPrintTimerToFile;
        j["repeatPrintTimerToStdOut"] = (unsigned long long)repeatPrintTimerToStdOut;
        j["runningDataFileDir"] = runningDataFileDir; j["runningDataFilePrefix"] = runningDataFilePrefix;
        j["runningDataFileAppendTime"] = runningDataFileAppendTime;
        j["doubleResolutionAtStep"] = (unsigned long long)doubleResolutionAtStep;
        return j;
    }
};

// CheckpointParams.hpp:26-80
struct CheckpointParams {
    bool startWithCheckpoint = false;
    std::string checkpointLoadFromDir = "notSet";
    int checkpointRepeat = 10;
    std::string checkpointWriteRootDir = ".";
    std::string checkpointWriteDirPrefix = "debug";
    bool checkpointWriteDirAppendTime = true;

    void getParamsFromJson(const json::Value& j) {
        parse_guard("CheckpointParams", [&] {
            startWithCheckpoint = j.at("startWithCheckpoint").get<bool>();
            checkpointLoadFromDir = j.at("checkpointLoadFromDir").get<std::string>();
            checkpointRepeat = j.at("checkpointRepeat").get<int>();
            checkpointWriteRootDir = j.at("checkpointWriteRootDir").get<std::string>();
            checkpointWriteDirPrefix = j.at("checkpointWriteDirPrefix").get<std::string>();
            checkpointWriteDirAppendTime = j.at("checkpointWriteDirAppendTime").get<bool>();
        });
    }
    json::Value getJson() const {
        json::Value j = json::Value::object();
        j["startWithCheckpoint"] = startWithCheckpoint; j["checkpointLoadFromDir"] = checkpointLoadFromDir;
        j["checkpointRepeat"] = checkpointRepeat; j["checkpointWriteRootDir"] = checkpointWriteRootDir;
        j["checkpointWriteDirPrefix"] = checkpointWriteDirPrefix;
        j["checkpointWriteDirAppendTime"] = checkpointWriteDirAppendTime;
        return j;
    }
};

// BinFileParams.hpp:26-106
struct BinFileParams {
    std::string filePath = "NoFilePath", name = "NoName", note = "NoNote", structName = "tDisk_colrow_Q4";
    tNi binFileSizeInStructs = 0;
    std::string coordsType = "uint16_t";
    bool hasGridtCoords = false, hasColRowtCoords = true;
    std::string reference = "absolute";
    tNi i0 = 0, j0 = 0, k0 = 0;
    std::string QDataType = "float";
    int QOutputLength = 4;

    void getParamsFromJson(const json::Value& j) {
        parse_guard("BinFileParams", [&] {
            filePath = j.at("filePath").get<std::string>();
            name = j.at("name").get<std::string>();
            note = j.at("note").get<std::string>();
            structName = j.at("structName").get<std::string>();
            binFileSizeInStructs = (tNi)j.at("binFileSizeInStructs").get<uint64_t>();
            coordsType = j.at("coordsType").get<std::string>();
            hasGridtCoords = j.at("hasGridtCoords").get<bool>();
            hasColRowtCoords = j.at("hasColRowtCoords").get<bool>();
            reference = j.at("reference").get<std::string>();
            i0 = (tNi)j.at("i0").get<uint64_t>();
            j0 = (tNi)j.at("j0").get<uint64_t>();
            k0 = (tNi)j.at("k0").get<uint64_t>();
            QDataType = j.at("QDataType").get<std::string>();
            QOutputLength = j.at("QOutputLength").get<int>();
        });
    }
    json::Value getJson() const {
        json::Value j = json::Value::object();
        j["filePath"] = filePath; j["name"] = name; j["note"] = note; j["structName"] = structName;
        j["binFileSizeInStructs"] = (long long)binFileSizeInStructs; j["coordsType"] = coordsType;
        j["hasGridtCoords"] = hasGridtCoords; j["hasColRowtCoords"] = hasColRowtCoords;
        j["reference"] = reference; j["i0"] = (long long)i0; j["j0"] = (long long)j0; j["k0"] = (long long)k0;
        j["QDataType"] = QDataType; j["QOutputLength"] = QOutputLength;
        return j;
    }
};

// ComputeUnitParams.hpp:26-123
struct ComputeUnitParams {
    int nodeID = 0, deviceID = 0, idi = 0, idj = 0, idk = 0;
    tNi x = 0, y = 0, z = 0, i0 = 0, j0 = 0, k0 = 0, ghost = 0, resolution = 0;
    std::string strQVecPrecision = "notSet", strQLength = "notSet", strMemoryLayout = "notSet";
    std::string strCollisonAlgo = "notSet", strStreamingAlgo = "notSet", strCompileFlag = "notSet";

    void getParamsFromJson(const json::Value& j) {
        parse_guard("ComputeUnitParams", [&] {
            nodeID = j.at("nodeID").get<int>();
            deviceID = j.at("deviceID").get<int>();
            idi = j.at("idi").get<int>();
            idj = j.at("idj").get<int>();
            idk = j.at("idk").get<int>();
            x = (tNi)j.at("x").get<uint64_t>();
            y = (tNi)j.at("y").get<uint64_t>();
            z = (tNi)j.at("z").get<uint64_t>();
            i0 = (tNi)j.at("i0").get<uint64_t>();
            j0 = (tNi)j.at("j0").get<uint64_t>();
            k0 = (tNi)j.at("k0").get<uint64_t>();
            ghost = (tNi)j.at("ghost").get<uint64_t>();
            resolution = (tNi)j.at("resolution").get<uint64_t>();
            strQVecPrecision = j.at("strQVecPrecision").get<std::string>();
            strQLength = j.at("strQLength").get<std::string>();
            strMemoryLayout = j.at("strMemoryLayout").get<std::string>();
            strCollisonAlgo = j.at("strCollisonAlgo").get<std::string>();
            strStreamingAlgo = j.at("strStreamingAlgo").get<std::string>();
            strCompileFlag = j.at("strCompileFlag").get<std::string>();
        });
    }
    json::Value getJson() const {
        json::Value j = json::Value::object();
        j["nodeID"] = nodeID; j["deviceID"] = deviceID; j["idi"] = idi; j["idj"] = idj; j["idk"] = idk;
        j["x"] = (long long)x; j["y"] = (long long)y; j["z"] = (long long)z;
        j["i0"] = (long long)i0; j["j0"] = (long long)j0; j["k0"] = (long long)k0;
        j["ghost"] = (long long)ghost; j["resolution"] = (long long)resolution;
        j["strQVecPrecision"] = strQVecPrecision; j["strQLength"] = strQLength;
        j["strMemoryLayout"] = strMemoryLayout; j["strCollisonAlgo"] = strCollisonAlgo;
        j["strStreamingAlgo"] = strStreamingAlgo; j["strCompileFlag"] = strCompileFlag;
        return j;
    }
};

inline json::Value read_json_file(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return json::Value::parse(ss.str());
}

}  // namespace tdlb
