// elastic_cube.cpp -- C++ driver of the explicit MPM substep through the
// MpmSimulationHip facade: a fixed-corotated cube (8 particles per cell on
// the sub-cell centres) dropping onto a sticky floor, as a scene written
// against the reference's MpmSimulationBase member names.
//
// usage: zmpm_elastic_cube <cells per side> <substeps> <fp32: 0|1> [frames] [restart dir]
//                          [--state FILE] [--autorestart MAX_DT]
// With frames > 0 the cube starts at 3 m/s downwards and the frame driver runs
// that many frames of 1/60 s with CFL substeps (calculateDt, TimeStepping::nextDt;
// max_dt = 1/100 s) instead of the fixed substeps.
// With --autorestart the frame driver instead runs simulate() with autorestart
// on, max_dt = MAX_DT and calculateDt() returning max_dt (SimulationBase's own
// calculateDt, SimulationBase.h:260-263): a max_dt that moves particles more
// than dx in a substep makes the device report the CFL condition, and simulate()
// rewinds to the last finished frame with max_dt halved until the substeps fit.
// With a restart directory (not "-") the fixed-substep run stops halfway, writes
// its state there (write(outputFileName())), and a second simulation built from
// the same scene setup restarts from that file and runs the other half.  The
// positions at the restart point also go to <restart dir>/partio_7.bgeo.
// --state FILE writes the final particle state (writeState) to FILE.
// prints one JSON line: particle count, centre of mass, mean velocity, substeps,
// time, restarts taken and the final max_dt.
#include "zmpm_simulation.hpp"

#include <cstdio>
#include <cstdlib>
#include <cstring>

// SimulationBase::calculateDt (SimulationBase.h:260-263): the frame driver's max_dt.
struct FixedDtSimulation : zmpm::MpmSimulationHip {
    double calculateDt() override { return max_dt; }
};

static void setup(zmpm::MpmSimulationHip& sim, int cells, bool fp32)
{
    sim.dx = 1.0 / 64;
    sim.gravity = { 0, -9.8, 0 };
    sim.precision = fp32 ? ZMPM_PREC_FP32 : ZMPM_PREC_FP64;
    sim.domain_nodes = 64;
    sim.autorestart = false;
    const double E = 1e4, nu = 0.3, rho = 1000;
    zmpm_material m {};
    m.model = ZMPM_MODEL_FIXED_COROTATED;
    m.mu = E / (2 * (1 + nu));                     // CorotatedElasticity.cpp:15-20
    m.lambda = E * nu / ((1 + nu) * (1 - 2 * nu));
    m.num_directors = 1;
    const int mat = sim.addMaterial(m);
    zmpm_collider floor {};
    floor.shape = ZMPM_SHAPE_HALF_SPACE;
    floor.type = ZMPM_STICKY;
    floor.origin[1] = 8 * sim.dx;
    floor.normal[1] = 1;
    sim.addCollisionObject(floor);
    const double h = sim.dx / 2, V0 = h * h * h;
    const double x0 = 16 * sim.dx, y0 = 8 * sim.dx, z0 = 16 * sim.dx;
    for (int i = 0; i < 2 * cells; i++)
        for (int j = 0; j < 2 * cells; j++)
            for (int k = 0; k < 2 * cells; k++)
                sim.addParticle({ x0 + (i + 0.5) * h, y0 + (j + 0.5) * h, z0 + (k + 0.5) * h }, rho * V0, V0, mat);
}

static const char* state_path = nullptr;

static int report(zmpm::MpmSimulationHip& sim, int substeps)
{
    if (state_path) {
        std::ofstream f(state_path, std::ios::binary);
        sim.writeState(f);
    }
    sim.sync();
    double cm[3] = { 0, 0, 0 }, vm[3] = { 0, 0, 0 }, M = 0;
    for (size_t p = 0; p < sim.X.size(); p++) {
        for (int a = 0; a < 3; a++) {
            cm[a] += sim.mass[p] * sim.X[p][a];
            vm[a] += sim.mass[p] * sim.V[p][a];
        }
        M += sim.mass[p];
    }
    std::printf("{\"n\": %zu, \"com\": [%.17g, %.17g, %.17g], \"mean_v\": [%.17g, %.17g, %.17g], \"substeps\": %d, "
                "\"time\": %.17g, \"restarts\": %d, \"max_dt\": %.17g}\n",
                sim.X.size(), cm[0] / M, cm[1] / M, cm[2] / M, vm[0] / M, vm[1] / M, vm[2] / M, substeps, sim.time,
                sim.restarts, sim.max_dt);
    return 0;
}

int main(int argc, char** argv)
{
    double autorestart_max_dt = 0;
    int nargs = 0;
    const char* pos[5] = { nullptr, nullptr, nullptr, nullptr, nullptr };
    for (int i = 1; i < argc; i++) {
        if (!std::strcmp(argv[i], "--state") && i + 1 < argc)
            state_path = argv[++i];
        else if (!std::strcmp(argv[i], "--autorestart") && i + 1 < argc)
            autorestart_max_dt = std::atof(argv[++i]);
        else if (nargs < 5)
            pos[nargs++] = argv[i];
    }
    const int cells = pos[0] ? std::atoi(pos[0]) : 16;
    const int steps = pos[1] ? std::atoi(pos[1]) : 20;
    const bool fp32 = pos[2] && std::atoi(pos[2]) != 0;
    const int frames = pos[3] ? std::atoi(pos[3]) : 0;
    const char* restart_dir = pos[4] && std::strcmp(pos[4], "-") ? pos[4] : nullptr;
    if (autorestart_max_dt > 0) {
        FixedDtSimulation fs;
        setup(fs, cells, fp32);
        for (auto& v : fs.V) v = { 0, -3, 0 };
        fs.frame_dt = 1.0 / 60;
        fs.max_dt = autorestart_max_dt;
        fs.max_dt_original = 1.0 / 60;
        fs.autorestart = true;
        fs.end_frame = frames > 0 ? frames : 1;
        fs.initialize();
        fs.simulate();
        return report(fs, 0);
    }
    zmpm::MpmSimulationHip sim;
    setup(sim, cells, fp32);
    int substeps = steps;
    if (frames > 0) {
        for (auto& v : sim.V) v = { 0, -3, 0 };
        sim.frame_dt = 1.0 / 60;
        sim.max_dt = 1.0 / 100;
        sim.initialize();
        substeps = 0;
        for (int f = 0; f < frames; f++) {
            sim.advanceOneFrame();
            substeps += sim.substeps_last_frame;
        }
    }
    else if (restart_dir) {
        sim.initialize();
        sim.advance(steps / 2, 1e-4);
        sim.frame = 7;
        sim.output_dir = restart_dir;
        sim.write_partio = true; // partio_7.bgeo beside the state file
        sim.write(sim.outputFileName());
        zmpm::MpmSimulationHip again;
        setup(again, cells, fp32);
        again.output_dir = restart_dir;
        again.restart(7);
        again.initialize();
        again.advance(steps - steps / 2, 1e-4);
        again.sync();
        return report(again, substeps);
    }
    else {
        sim.initialize();
        sim.advance(steps, 1e-4);
    }
    return report(sim, substeps);
}
