// tdlb_run -- the reference's main loop (tdlbcpp/src/main.cpp:55-555) on one MI355X.
//
//   tdlb_run -i input.json [-c checkpoint_dir] [--layout IJKL|LIJK] [--precision f32|f64]
//            [--per-call] [--steps N] [--print-params] [--nu-from-json] [--write-planes] [--no-output]
//
// Reads the reference's input JSON, builds one ComputeUnit (ghost = multiStep),
// initialises it or loads a checkpoint, generates the Rushton tank, sets nu from
// calcNuAndRe_m, runs the initial forcing and then num_steps iterations of
// collision -> bounceBack -> streamingPush -> moments -> forcing, writing
// checkpoints every checkpointRepeat steps and, as main.cpp:462-487 does after
// every step, the vorticity pictures of OutputParams' XY_vorticity_planes and
// XZ_vorticity_planes that are due.  --write-planes also writes the
// savePlaneXZ records of XZ_planes (main.cpp:478-482 has that call commented
// out); --no-output writes no plots.  The default path runs the fused device
// loop (tdlb_cu_run); --per-call drives the five reference calls per step with
// host-generated IB lists, as main.cpp does.
// --print-params parses the input and prints the parsed parameters (no GPU).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "ComputeUnit.hpp"

using namespace tdlb;

namespace {

void usage() {
    std::fprintf(stderr,
                 "usage: tdlb_run -i input.json [-c checkpoint_dir] [--layout IJKL|LIJK] [--precision f32|f64]\n"
                 "                [--per-call] [--steps N] [--print-params] [--nu-from-json] [--write-planes]\n"
                 "                [--no-output]\n");
}

std::vector<tdlb_ib_point> generate(const tdlb_rushton_params& rp, int group, double angle, double delta) {
    int64_t n = 0;
    throw_if(tdlb_rushton_generate(&rp, group, angle, delta, nullptr, 0, &n), "geometry");
    std::vector<tdlb_ib_point> v(n);
    throw_if(tdlb_rushton_generate(&rp, group, angle, delta, v.data(), n, &n), "geometry");
    return v;
}

// an entry of OutputParams' plane lists (OrthoPlaneParams,
// OrthoPlaneVorticityParams): the fields the main loop reads
struct PlaneOutput {
    tNi cutAt = 0;
    tStep repeat = 0, start_at_step = 0;
    int jpegCompression = 100;
    // main.cpp:470-485: repeat && step >= start_at_step && (step - start) % repeat == 0
    bool due(tStep step) const {
        return repeat && step >= start_at_step && (step - start_at_step) % repeat == 0;
    }
};

std::vector<PlaneOutput> plane_list(const json::Value& output, const char* key) {
    std::vector<PlaneOutput> out;
    if (!output.contains(key)) return out;
    for (const json::Value& e : output.at(key).elements()) {
        PlaneOutput p;
        p.cutAt = e.value<tNi>("cutAt", 0);
        p.repeat = e.value<tStep>("repeat", 0);
        p.start_at_step = e.value<tStep>("start_at_step", 0);
        p.jpegCompression = e.value<int>("jpegCompression", 100);
        out.push_back(p);
    }
    return out;
}

std::vector<PosPolar> to_pos(const std::vector<tdlb_ib_point>& v) {
    std::vector<PosPolar> out(v.size());
    for (size_t n = 0; n < v.size(); n++)
        out[n] = {v[n].i, v[n].j, v[n].k, v[n].i_frac, v[n].k_frac, v[n].u_delta, v[n].v_delta, v[n].w_delta};
    return out;
}

}  // namespace

int main(int argc, char** argv) {
    std::string input = "input_debug_gridx60_numSteps20.json", ckdir;
    int layout = TDLB_LAYOUT_IJKL, precision = TDLB_PREC_F32;
    bool per_call = false, print_only = false, nu_from_json = false, write_planes = false, no_output = false;
    long long steps_override = -1;
    for (int a = 1; a < argc; a++) {
        std::string s = argv[a];
        auto next = [&]() -> std::string {
            if (a + 1 >= argc) { usage(); std::exit(1); }
            return argv[++a];
        };
        if (s == "-i" || s == "--input_file") input = next();
        else if (s == "-c" || s == "--checkpoint_dir") ckdir = next();
        else if (s == "--layout") layout = next() == "LIJK" ? TDLB_LAYOUT_LIJK : TDLB_LAYOUT_IJKL;
        else if (s == "--precision") precision = next() == "f64" ? TDLB_PREC_F64 : TDLB_PREC_F32;
        else if (s == "--per-call") per_call = true;
        else if (s == "--steps") steps_override = std::atoll(next().c_str());
        else if (s == "--print-params") print_only = true;
        else if (s == "--nu-from-json") nu_from_json = true;
        else if (s == "--write-planes") write_planes = true;
        else if (s == "--no-output") no_output = true;
        else if (s == "-h" || s == "--help") { usage(); return 1; }
        else { usage(); return 1; }
    }
    if (!ckdir.empty()) input = ckdir + "/AllParams.0.0.0.device.json";  // main.cpp:131

    GridParams grid;
    FlowParams flow;
    RunningParams running;
    CheckpointParams checkpoint;
    BinFileParams binFormat;
    json::Value output = json::Value::object();
    try {
        json::Value j = read_json_file(input);
        grid.getParamsFromJson(j.at("GridParams"));
        flow.getParamsFromJson(j.at("FlowParams"));
        running.getParamsFromJson(j.at("RunningParams"));
        checkpoint.getParamsFromJson(j.at("CheckpointParams"));
        binFormat.getParamsFromJson(j.at("BinFileParams"));
        output = j.at("OutputParams");
        if (!output.contains("outputRootDir")) throw std::runtime_error("Failed to parse OutputParams: outputRootDir");
    } catch (std::exception& e) {
        std::cerr << "Exception reached parsing input json: " << e.what() << std::endl;
        return EXIT_FAILURE;
    }
    if (steps_override >= 0) running.num_steps = running.step + (tStep)steps_override - 1;

    tdlb_rushton_params rp{};
    if (tdlb_rushton_defaults(grid.x, grid.y, flow.uav, (int64_t)running.impellerStartupStepsUntilNormalSpeed, &rp)) {
        std::cerr << tdlb_last_error() << std::endl;
        return EXIT_FAILURE;
    }
    // RushtonTurbine(grid.x); flow.calcNuAndRe_m(outerRadius)  (main.cpp:274-276).
    // main.cpp applies it to its own copy of FlowParams after the ComputeUnit was
    // constructed, so the reference's collision runs with the JSON "nu"; here the
    // derived viscosity reaches the ComputeUnit unless --nu-from-json is given.
    if (!nu_from_json) {
        tdlb_flow cf = ComputeUnit::to_c(flow);
        throw_if(tdlb_flow_calc_nu_re_m(&cf, (int32_t)rp.blade_outer_radius, precision), "calcNuAndRe_m");
        flow.nu = cf.nu;
        flow.Re_m = cf.re_m;
    }

    ComputeUnitParams cu;
    cu.nodeID = 0; cu.deviceID = 0; cu.idi = cu.idj = cu.idk = 0;
    cu.x = grid.x; cu.y = grid.y; cu.z = grid.z; cu.i0 = cu.j0 = cu.k0 = 0;
    cu.ghost = grid.multiStep; cu.resolution = 1;
    cu.strCompileFlag = "WITH_GPU_HIP";
    cu.strCollisonAlgo = flow.collision;
    cu.strStreamingAlgo = flow.streaming;
    cu.strQVecPrecision = precision == TDLB_PREC_F32 ? "float" : "double";
    cu.strQLength = "18";
    cu.strMemoryLayout = layout == TDLB_LAYOUT_IJKL ? "IJKL" : "LIJK";

    if (print_only) {
        json::Value j = json::Value::object();
        j["GridParams"] = grid.getJson();
        j["FlowParams"] = flow.getJson();
        j["RunningParams"] = running.getJson();
        j["CheckpointParams"] = checkpoint.getJson();
        j["BinFileParams"] = binFormat.getJson();
        j["OutputParams"] = output;
        j["ComputeUnitParams"] = cu.getJson();
        std::cout << j.dump(4) << std::endl;
        return 0;
    }

    try {
        DiskOutputTree tree(output, checkpoint);
        tree.setParams(cu, grid, flow.getJson(), running, output, checkpoint);
        ComputeUnit lb(cu, flow, tree, layout, precision);
        if (ckdir.empty()) {
            lb.initialise(flow.initialRho);
        } else {
            running = lb.checkpoint_read(ckdir, "device");
            lb.setFlow(flow);
        }
        std::cout << "ComputeUnit " << lb.x << "x" << lb.y << "x" << lb.z << " streaming=" << flow.streaming
                  << " layout=" << cu.strMemoryLayout << " nu=" << flow.nu << " LES=" << flow.useLES << std::endl;

        std::vector<PosPolar> geomFORCING;  // per-call: the lists of the last forcing pass
        const std::vector<PlaneOutput> xy_vort = plane_list(output, "XY_vorticity_planes");
        const std::vector<PlaneOutput> xz_vort = plane_list(output, "XZ_vorticity_planes");
        const std::vector<PlaneOutput> xz_planes = plane_list(output, "XZ_planes");
        // main.cpp:462-487 after each step: exclusion from the forcing geometry,
        // then the vorticity pictures (and, on request, the XZ plane records) due
        auto write_output = [&]() {
            if (no_output) return;
            bool any = false;
            for (auto* l : {&xy_vort, &xz_vort}) for (const PlaneOutput& p : *l) any |= p.due(running.step);
            if (write_planes) for (const PlaneOutput& p : xz_planes) any |= p.due(running.step);
            if (!any) return;
            if (per_call) lb.setOutputExcludePoints(rp, geomFORCING);
            else lb.setOutputExcludePoints();
            for (const PlaneOutput& p : xy_vort)
                if (p.due(running.step)) lb.calcVorticityXY(p.cutAt, running, p.jpegCompression);
            if (write_planes)
                for (const PlaneOutput& p : xz_planes)
                    if (p.due(running.step)) lb.savePlaneXZ(p.cutAt, binFormat, running);
            for (const PlaneOutput& p : xz_vort)
                if (p.due(running.step)) lb.calcVorticityXZ(p.cutAt, running, p.jpegCompression);
            lb.unsetOutputExcludePoints();
        };

        const bool rnd32 = precision == TDLB_PREC_F32;
        auto inc = [&](tStep s) {
            double d = 0;
            tdlb_rushton_step_increment(&rp, s, &d);
            return rnd32 ? (double)(float)d : d;
        };
        std::vector<PosPolar> fixed, nonupd;
        auto t0 = std::chrono::steady_clock::now();
        const tStep first = running.step;
        if (!per_call) {
            throw_if(tdlb_cu_bind_rushton(lb.handle(), &rp, running.step, running.angle, flow.alpha, flow.beta),
                     "bind_rushton");
        } else {
            const double d0 = inc(running.step);
            fixed = to_pos(generate(rp, TDLB_GEOM_FIXED, 0, 0));
            nonupd = to_pos(generate(rp, TDLB_GEOM_ROTATING_NON_UPDATING, 0, d0));
            std::vector<PosPolar> all = fixed;
            all.insert(all.end(), nonupd.begin(), nonupd.end());
            auto rot = to_pos(generate(rp, TDLB_GEOM_ROTATING, running.angle, d0));
            all.insert(all.end(), rot.begin(), rot.end());
            lb.forcing(all, flow.alpha, flow.beta);
        }
        for (tStep step = first; step <= running.num_steps; step++) {
            if (running.step + 1 == running.doubleResolutionAtStep) {
                // main.cpp:349-399, at the top of the step that doubles: regrid, the
                // tank of the new size, angle += its increment, forcing on the new lists
                lb.doubleResolutionFullCU();
                cu.x = lb.x; cu.y = lb.y; cu.z = lb.z; cu.resolution = lb.resolution;
                grid.x = lb.x; grid.y = lb.y; grid.z = lb.z;
                throw_if(tdlb_rushton_defaults(lb.x, lb.y, flow.uav, (int64_t)running.impellerStartupStepsUntilNormalSpeed,
                                               &rp),
                         "rushton_defaults");
                if (!nu_from_json) {
                    tdlb_flow cf = ComputeUnit::to_c(flow);
                    throw_if(tdlb_flow_calc_nu_re_m(&cf, (int32_t)rp.blade_outer_radius, precision), "calcNuAndRe_m");
                    flow.nu = cf.nu;
                    flow.Re_m = cf.re_m;
                    lb.setFlow(flow);
                }
                tree.setParams(cu, grid, flow.getJson(), running, output, checkpoint);
                const tStep sd = running.step + 1;
                const double d0 = inc(sd);
                running.angle += d0;
                if (!per_call) {
                    throw_if(tdlb_cu_bind_rushton(lb.handle(), &rp, sd, running.angle, flow.alpha, flow.beta),
                             "bind_rushton");
                } else {
                    fixed = to_pos(generate(rp, TDLB_GEOM_FIXED, 0, 0));
                    nonupd = to_pos(generate(rp, TDLB_GEOM_ROTATING_NON_UPDATING, 0, d0));
                    std::vector<PosPolar> all = fixed;
                    all.insert(all.end(), nonupd.begin(), nonupd.end());
                    auto rot = to_pos(generate(rp, TDLB_GEOM_ROTATING, running.angle, d0));
                    all.insert(all.end(), rot.begin(), rot.end());
                    lb.forcing(all, flow.alpha, flow.beta);
                }
                std::printf("INCREASING RESOLUTION TO grid.x %lld at step %lu\n", (long long)lb.x,
                            (unsigned long)sd);
            }
            if (!per_call) {
                uint64_t s = running.step;
                double a = running.angle;
                throw_if(tdlb_cu_run(lb.handle(), 1, &s, &a), "run");
                running.step = s;
                running.angle = a;
            } else {
                running.incrementStep();
                const double d = inc(running.step);
                running.angle += d;
                std::vector<PosPolar> all = fixed;
                all.insert(all.end(), nonupd.begin(), nonupd.end());
                auto rot = to_pos(generate(rp, TDLB_GEOM_ROTATING, running.angle, d));
                all.insert(all.end(), rot.begin(), rot.end());
                lb.collision();
                lb.bounceBackBoundary();
                lb.streamingPush();
                lb.moments();
                lb.forcing(all, flow.alpha, flow.beta);
                geomFORCING.swap(all);
            }
            write_output();
            if (checkpoint.checkpointRepeat && (running.step % checkpoint.checkpointRepeat == 0))
                lb.checkpoint_write("device", running, binFormat);
            if (step == first || step == running.num_steps || (step % 10) == 0) {
                const double revs = running.angle * ((180.0 / M_PI) / 360);
                std::printf("Node 0 Step %lu/%lu, Angle: %.4g (%.1f revs)\n", (unsigned long)step,
                            (unsigned long)running.num_steps, running.angle, revs);
            }
        }
        throw_if(tdlb_cu_synchronize(lb.handle()), "synchronize");
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        const double n = (double)(running.num_steps + 1 - first);
        std::printf("%.0f steps in %.3f s: %.1f MLUPS\n", n, s, (double)lb.x * lb.y * lb.z * n / s / 1e6);
        if (!lb.lastCheckpointDir.empty()) std::printf("last checkpoint: %s\n", lb.lastCheckpointDir.c_str());
    } catch (std::exception& e) {
        std::cerr << "tdlb_run: " << e.what() << std::endl;
        return EXIT_FAILURE;
    }
    return 0;
}
