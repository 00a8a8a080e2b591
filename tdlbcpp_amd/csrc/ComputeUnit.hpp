// C++ host-side ComputeUnit over the C ABI: the reference's ComputeUnitBase
// surface (tdlbcpp/src/ComputeUnit.h:39-204) backed by a handle in HBM, plus the
// checkpoint and plot naming of DiskOutputTree (tdlbcpp/src/DiskOutputTree.h:
// 331-410, 418-590) and the plane and vorticity outputs of
// ComputeUnitOutput.hpp:113-245, 354-447.
//
// The reference selects layout, streaming and precision with template
// arguments; here they are constructor arguments because the device code is
// compiled for all of them.  Errors: the C ABI's codes become
// std::runtime_error (the reference aborts through checkCudaErrors or exits).
#pragma once

#include <sys/stat.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <ctime>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tdlb.h"
#include "jpeg_gray.hpp"
#include "params.hpp"

namespace tdlb {

// The PosPolar<tNi,T> fields that forcing reads (tdlbcpp/src/Forcing.hpp:76-128).
struct PosPolar {
    tNi i, j, k;
    double iCartFraction, kCartFraction, uDelta, vDelta, wDelta;
};

inline void throw_if(int rc, const char* what) {
    if (rc != TDLB_OK) throw std::runtime_error(std::string(what) + ": " + tdlb_last_error());
}

// Directory and file naming of checkpoints and the AllParams JSON.
class DiskOutputTree {
public:
    DiskOutputTree(const json::Value& output, const CheckpointParams& c) : output_(output), ckpt_(c) {
        char buf[80];
        std::time_t t = std::time(nullptr);
        std::strftime(buf, sizeof buf, "%Y_%m_%d_%H_%M_%S", std::localtime(&t));
        initTime = buf;
        checkpointWriteDir = ckpt_.checkpointWriteRootDir + "/" + checkpointName();
    }
    void setParams(const ComputeUnitParams& cu, const GridParams& g, const json::Value& flow, const RunningParams& r,
                   const json::Value& output, const CheckpointParams& c) {
        cu_ = cu;
        grid_ = g;
        flow_ = flow;
        running_ = r;
        output_ = output;
        ckpt_ = c;
        checkpointWriteDir = ckpt_.checkpointWriteRootDir + "/" + checkpointName();
    }
    std::string formatId() const {
        return std::to_string(cu_.idi) + "." + std::to_string(cu_.idj) + "." + std::to_string(cu_.idk);
    }
    std::string checkpointName() const {
        std::string name;
        if (!ckpt_.checkpointWriteDirPrefix.empty()) name += ckpt_.checkpointWriteDirPrefix + "_";
        name += "checkpoint";
        if (ckpt_.checkpointWriteDirAppendTime) name += "_" + initTime;
        return name;
    }
    std::string getCheckpointDirName(const RunningParams& run, bool create = true) const {
        const std::string dir = checkpointWriteDir + "/" + checkpointName() + "_step_" + std::to_string(run.step);
        if (create) {
            ::mkdir(ckpt_.checkpointWriteRootDir.c_str(), 0775);
            ::mkdir(checkpointWriteDir.c_str(), 0775);
            ::mkdir(dir.c_str(), 0775);
        }
        return dir;
    }
    std::string getAllParamsFilePath(const std::string& dir, const std::string& unit) const {
        return dir + "/AllParams." + formatId() + "." + unit;
    }

    // ---- plot naming (DiskOutputTree.h:331-410): <outputRootDir>_<initTime>/
    // <prefix>.<plotType>.V5.step_<8-digit step>[.cut_<n>], files CUid.<id>.V5.*
    std::string outputDir() const {
        std::string root = "Error_UnInitialised_Output_Dir";
        if (output_.contains("outputRootDir")) root = output_.at("outputRootDir").get<std::string>();
        return root + "_" + initTime;
    }
    static std::string formatStep(tStep step) {
        char buf[32];
        std::snprintf(buf, sizeof buf, "%08lu", (unsigned long)step);
        return buf;
    }
    std::string formatPlotDir(const std::string& prefix, const std::string& plotType, tStep step) const {
        return outputDir() + "/" + prefix + "." + plotType + ".V5.step_" + formatStep(step);
    }
    std::string formatXYPlaneDir(tStep step, tNi atK, const std::string& prefix = "plot") const {
        return formatPlotDir(prefix, "xyPlane", step) + ".cut_" + std::to_string(atK);
    }
    std::string formatXZPlaneDir(tStep step, tNi atJ, const std::string& prefix = "plot") const {
        return formatPlotDir(prefix, "xzPlane", step) + ".cut_" + std::to_string(atJ);
    }
    std::string formatCUid() const { return "CUid." + formatId(); }
    std::string formatQVecBinFileNamePath(const std::string& dir) const {
        return dir + "/" + formatCUid() + ".V5.QVec.bin";
    }
    std::string formatF3BinFileNamePath(const std::string& dir) const {
        return dir + "/" + formatCUid() + ".V5.F3.bin";
    }
    // <bin.filePath>.json holding every parameter block (writeAllParamsJson)
    void writeAllParamsJson(const BinFileParams& bin, const RunningParams& run) const {
        const std::string path = bin.filePath + ".json";
        FILE* fp = std::fopen(path.c_str(), "w");
        if (!fp) throw std::runtime_error("cannot write " + path);
        const std::string text = getJsonParams(bin, run).dump();
        const bool ok = std::fwrite(text.data(), 1, text.size(), fp) == text.size();
        if (std::fclose(fp) != 0 || !ok) throw std::runtime_error("cannot write " + path);
    }
    // mkdir -p
    static void createDir(const std::string& path) {
        for (size_t p = path.find('/', 1); ; p = path.find('/', p + 1)) {
            const std::string d = p == std::string::npos ? path : path.substr(0, p);
            if (!d.empty()) ::mkdir(d.c_str(), 0775);
            if (p == std::string::npos) break;
        }
    }
    json::Value getJsonParams(const BinFileParams& bin, const RunningParams& run) const {
        json::Value j = json::Value::object();
        j["ComputeUnitParams"] = cu_.getJson();
        j["BinFileParams"] = bin.getJson();
        j["GridParams"] = grid_.getJson();
        j["FlowParams"] = flow_;
        j["RunningParams"] = run.getJson();
        j["OutputParams"] = output_;
        j["CheckpointParams"] = ckpt_.getJson();
        return j;
    }

    std::string initTime;
    std::string checkpointWriteDir;

private:
    ComputeUnitParams cu_;
    GridParams grid_;
    json::Value flow_ = json::Value::object();
    RunningParams running_;
    json::Value output_ = json::Value::object();
    CheckpointParams ckpt_;
};

class ComputeUnit {
public:
    // grid geometry (ComputeUnit.h:44-65, ComputeUnit.hpp:25-40)
    int idi = 0, idj = 0, idk = 0, nodeID = 0, deviceID = 0;
    tNi resolution = 0, x = 0, y = 0, z = 0, i0 = 0, j0 = 0, k0 = 0;
    tNi xg = 0, yg = 0, zg = 0, xg0 = 0, yg0 = 0, zg0 = 0, xg1 = 0, yg1 = 0, zg1 = 0, ghost = 0;
    size_t size = 0;
    FlowParams flow;
    int layout = TDLB_LAYOUT_IJKL, streaming = TDLB_STREAM_SIMPLE, precision = TDLB_PREC_F32;

    ComputeUnit(const ComputeUnitParams& cu, const FlowParams& f, DiskOutputTree& tree, int layout_,
                int precision_ = TDLB_PREC_F32)
        : flow(f), layout(layout_), precision(precision_), tree_(&tree) {
        streaming = f.streaming == "Simple" ? TDLB_STREAM_SIMPLE : TDLB_STREAM_ESOTWIST;
        init(cu);
    }
    ComputeUnit(const ComputeUnit&) = delete;
    ComputeUnit& operator=(const ComputeUnit&) = delete;
    ~ComputeUnit() { tdlb_cu_destroy(h_); }

    tdlb_cu_t handle() const { return h_; }

    // ---- operator surface
    void initialise(double rho) { throw_if(tdlb_cu_initialise(h_, rho), "initialise"); }
    void collision() { throw_if(tdlb_cu_collision(h_), "collision"); }
    void moments() { throw_if(tdlb_cu_moments(h_), "moments"); }
    void streamingPush() { throw_if(tdlb_cu_stream_push(h_), "streamingPush"); }
    void streamingPull() { throw_if(tdlb_cu_stream_pull(h_), "streamingPull"); }
    void bounceBackBoundary() { throw_if(tdlb_cu_bounce_back(h_), "bounceBackBoundary"); }
    void forcing(const std::vector<PosPolar>& geom, double alfa, double beta) {
        std::vector<tdlb_ib_point> p(geom.size());
        for (size_t n = 0; n < geom.size(); n++)
            p[n] = {geom[n].i, geom[n].j, geom[n].k, geom[n].iCartFraction, geom[n].kCartFraction, geom[n].uDelta,
                    geom[n].vDelta, geom[n].wDelta};
        throw_if(tdlb_cu_forcing(h_, p.data(), (int64_t)p.size(), alfa, beta), "forcing");
    }
    bool evenStep() const {
        int e = 1;
        throw_if(tdlb_cu_get_even_step(h_, &e), "evenStep");
        return e != 0;
    }
    void setFlow(const FlowParams& f) {
        flow = f;
        tdlb_flow cf = to_c(f);
        throw_if(tdlb_cu_set_flow(h_, &cf), "setFlow");
    }
    // doubleResolutionFullCU (ComputeUnit.hpp:257-316): sizes double, resolution += 2
    void doubleResolutionFullCU() {
        throw_if(tdlb_cu_double_resolution(h_), "doubleResolutionFullCU");
        x *= 2; y *= 2; z *= 2;
        resolution += 2;
        xg = x + 2 * ghost; yg = y + 2 * ghost; zg = z + 2 * ghost;
        xg0 = xg - 1; yg0 = yg - 1; zg0 = zg - 1;
        xg1 = xg - 2; yg1 = yg - 2; zg1 = zg - 2;
        size = (size_t)xg * yg * zg;
    }

    // ---- checkpoint (ComputeUnit.hpp:322-439), JSON and raw arrays byte-compatible.
    // The AllParams JSON carries one extension key, "ComputeUnitState": {"evenStep"},
    // which the reference's reader ignores (it does not save evenStep).
    void checkpoint_write(const std::string& unit_name, const RunningParams& run, const BinFileParams& bin_in) {
        const std::string dir = tree_->getCheckpointDirName(run);
        BinFileParams bin = bin_in;
        bin.filePath = tree_->getAllParamsFilePath(dir, unit_name);
        bin.structName = "checkpoint";
        bin.binFileSizeInStructs = (tNi)size;
        bin.coordsType = "none";
        bin.hasGridtCoords = false;
        bin.hasColRowtCoords = false;
        bin.QDataType = "none";
        bin.QOutputLength = 18;
        json::Value j = tree_->getJsonParams(bin, run);
        json::Value st = json::Value::object();
        st["evenStep"] = evenStep();
        j["ComputeUnitState"] = st;
        const std::string text = j.dump();
        throw_if(tdlb_cu_checkpoint_write(h_, dir.c_str(), unit_name.c_str(), text.c_str()), "checkpoint_write");
        lastCheckpointDir = dir;
    }

    // Reads AllParams, re-initialises the ComputeUnit from its ComputeUnitParams
    // (sizes may change) and loads the arrays; returns the RunningParams stored.
    RunningParams checkpoint_read(const std::string& dirname, const std::string& unit_name) {
        const std::string path = tree_->getAllParamsFilePath(dirname, unit_name) + ".json";
        json::Value j = read_json_file(path);
        ComputeUnitParams cu;
        cu.getParamsFromJson(j.at("ComputeUnitParams"));
        FlowParams f;
        f.getParamsFromJson(j.at("FlowParams"));
        RunningParams run;
        run.getParamsFromJson(j.at("RunningParams"));
        flow = f;
        streaming = f.streaming == "Simple" ? TDLB_STREAM_SIMPLE : TDLB_STREAM_ESOTWIST;
        tdlb_cu_destroy(h_);
        h_ = nullptr;
        init(cu);
        int missing = 0;
        throw_if(tdlb_cu_checkpoint_read(h_, dirname.c_str(), unit_name.c_str(), &missing), "checkpoint_read");
        if (j.contains("ComputeUnitState") && j.at("ComputeUnitState").contains("evenStep"))
            throw_if(tdlb_cu_set_even_step(h_, j.at("ComputeUnitState").at("evenStep").get<bool>() ? 1 : 0),
                     "checkpoint_read");
        missingArrays = missing;
        return run;
    }

    std::string lastCheckpointDir;
    int missingArrays = 0;

    // ---- outputs (ComputeUnitOutput.hpp, main.cpp:466-487)
    // ExcludeOutputPoints: outside the tank and every cell holding a point of
    // the bound geometry (setOutputExcludePoints(geomFORCING))
    std::vector<uint8_t> ExcludeOutputPoints;
    void setOutputExcludePoints() {
        ExcludeOutputPoints.assign(size, 0);
        throw_if(tdlb_cu_output_exclude(h_, ExcludeOutputPoints.data(), size), "setOutputExcludePoints");
    }
    // the same from host lists (a ComputeUnit driven by forcing() calls, with
    // no geometry bound): the tank's solid mask (main.cpp:309-315), then every
    // point of geom at indexPlusGhost (ComputeUnitOutput.hpp:29-34)
    void setOutputExcludePoints(const tdlb_rushton_params& rp, const std::vector<PosPolar>& geom) {
        ExcludeOutputPoints.assign(size, 0);
        throw_if(tdlb_rushton_solid_mask(&rp, xg, yg, zg, ghost, ExcludeOutputPoints.data()),
                 "setOutputExcludePoints");
        for (const PosPolar& p : geom)
            ExcludeOutputPoints[((size_t)(p.i + ghost) * yg + (size_t)(p.j + ghost)) * zg + (size_t)(p.k + ghost)] = 1;
    }
    void unsetOutputExcludePoints() { ExcludeOutputPoints.clear(); }

    // savePlaneXZ<float, 4> (ComputeUnitOutput.hpp:354-447): tDiskGrid<float,4>
    // records of the plane j = cutAt and tDiskGrid<float,3> records of its
    // non-zero forces, each file with its AllParams JSON.  The 2 alignment
    // bytes after the coordinates of every record are written as zero (the
    // reference leaves them unset).
    void savePlaneXZ(tNi cutAt, BinFileParams bin, const RunningParams& run) {
        const size_t cells = (size_t)xg1 * zg1;
        std::vector<uint8_t> qv(cells * 24, 0), f3(cells * 20, 0);
        int64_t nq = 0, nf = 0;
        throw_if(tdlb_cu_plane_xz(h_, cutAt, exclude(), qv.data(), &nq, f3.data(), &nf), "savePlaneXZ");
        const std::string dir = tree_->formatXZPlaneDir(run.step, cutAt);
        DiskOutputTree::createDir(dir);
        bin.filePath = tree_->formatQVecBinFileNamePath(dir);
        bin.binFileSizeInStructs = (tNi)nq;
        std::cout << "Writing output to: " << bin.filePath << std::endl;
        tree_->writeAllParamsJson(bin, run);
        write_file(bin.filePath, qv.data(), (size_t)nq * 24);
        bin.QOutputLength = 3;
        bin.binFileSizeInStructs = (tNi)nf;
        bin.filePath = tree_->formatF3BinFileNamePath(dir);
        tree_->writeAllParamsJson(bin, run);
        write_file(bin.filePath, f3.data(), (size_t)nf * 20);
    }

    // calcVorticityXZ / calcVorticityXY (ComputeUnitOutput.hpp:113-245): the
    // picture floor(255 (v - min) / (max - min)) of v = log|curl u|^2 on the
    // plane, min and max over its finite values, as a JPEG of quality 100 named
    // by the plane's plot directory.  Excluded and non-finite cells are 0 (the
    // reference leaves excluded cells unset and converts -inf out of range).
    void calcVorticityXZ(tNi j, const RunningParams& run, int /*jpegCompression: quality 100 as the reference*/) {
        vorticity_picture(0, j, tree_->formatXZPlaneDir(run.step, j) + ".jpeg");
    }
    void calcVorticityXY(tNi k, const RunningParams& run, int /*jpegCompression*/) {
        vorticity_picture(1, k, tree_->formatXYPlaneDir(run.step, k) + ".jpeg");
    }

    static tdlb_flow to_c(const FlowParams& f) {
        tdlb_flow c{};
        c.initial_rho = f.initialRho; c.re_m_nondimensional = f.reMNonDimensional; c.uav = f.uav;
        c.cs0 = f.cs0; c.g3 = f.g3; c.nu = f.nu; c.fx0 = f.fx0; c.fy0 = f.fy0; c.fz0 = f.fz0;
        c.re_m = f.Re_m; c.re_f = f.Re_f; c.uf = f.uf; c.alpha = f.alpha; c.beta = f.beta;
        c.use_les = f.useLES ? 1 : 0;
        c.collision = f.collision == "EgglesSomers" ? 0 : 1;
        return c;
    }

private:
    tdlb_cu_t h_ = nullptr;
    DiskOutputTree* tree_;

    const uint8_t* exclude() const { return ExcludeOutputPoints.empty() ? nullptr : ExcludeOutputPoints.data(); }

    static void write_file(const std::string& path, const void* data, size_t bytes) {
        FILE* fp = std::fopen(path.c_str(), "wb");
        if (!fp) throw std::runtime_error("cannot write " + path);
        const bool ok = bytes == 0 || std::fwrite(data, 1, bytes, fp) == bytes;
        if (std::fclose(fp) != 0 || !ok) throw std::runtime_error("cannot write " + path);
    }

    template <typename T>
    std::vector<uint8_t> picture(int axis, tNi cut) const {
        const tNi n2 = axis == 0 ? zg1 : yg1;
        std::vector<T> v((size_t)xg1 * n2);
        throw_if(tdlb_cu_vorticity(h_, axis, cut, exclude(), v.data(), v.size() * sizeof(T)), "vorticity");
        bool have = false;
        T lo = 0, hi = 0;
        for (T a : v)
            if (std::isfinite(a)) {
                if (!have || a < lo) lo = a;
                if (!have || a > hi) hi = a;
                have = true;
            }
        // XZ: pict[zg1 (i-1) + (k-1)] of a zg1-row, xg1-wide picture; XY:
        // pict[(i-1) + xg1 (j-1)] of a yg1-row picture
        std::vector<uint8_t> pict((size_t)xg1 * n2, 0);
        for (tNi i = 0; i < xg1; i++)
            for (tNi b = 0; b < n2; b++) {
                const T a = v[(size_t)i * n2 + b];
                uint8_t px = 0;
                if (std::isfinite(a)) {
                    const T s = std::floor(T(255) * ((a - lo) / (hi - lo)));
                    if (std::isfinite(s)) px = (uint8_t)(s < T(0) ? T(0) : s > T(255) ? T(255) : s);
                }
                pict[axis == 0 ? (size_t)i * n2 + b : (size_t)i + (size_t)xg1 * b] = px;
            }
        return pict;
    }

    void vorticity_picture(int axis, tNi cut, const std::string& path) const {
        const std::vector<uint8_t> pict = precision == TDLB_PREC_F64 ? picture<double>(axis, cut)
                                                                     : picture<float>(axis, cut);
        DiskOutputTree::createDir(tree_->outputDir());
        if (!jpeg::write_gray(path, pict.data(), (int)xg1, (int)(axis == 0 ? zg1 : yg1), 100))
            throw std::runtime_error("cannot write " + path);
    }

    void init(const ComputeUnitParams& cu) {
        idi = cu.idi; idj = cu.idj; idk = cu.idk; nodeID = cu.nodeID; deviceID = cu.deviceID;
        resolution = cu.resolution; x = cu.x; y = cu.y; z = cu.z; i0 = cu.i0; j0 = cu.j0; k0 = cu.k0;
        ghost = cu.ghost;
        xg = x + 2 * ghost; yg = y + 2 * ghost; zg = z + 2 * ghost;
        xg0 = xg - 1; yg0 = yg - 1; zg0 = zg - 1;
        xg1 = xg - 2; yg1 = yg - 2; zg1 = zg - 2;
        size = (size_t)xg * yg * zg;
        tdlb_cu_params p{};
        p.x = x; p.y = y; p.z = z; p.ghost = ghost; p.i0 = i0; p.j0 = j0; p.k0 = k0; p.resolution = resolution;
        p.idi = idi; p.idj = idj; p.idk = idk; p.node_id = nodeID; p.device_id = deviceID < 0 ? 0 : deviceID;
        p.layout = layout; p.streaming = streaming; p.precision = precision;
        tdlb_flow cf = to_c(flow);
        throw_if(tdlb_cu_create(&p, &cf, &h_), "ComputeUnit");
    }
};

}  // namespace tdlb
