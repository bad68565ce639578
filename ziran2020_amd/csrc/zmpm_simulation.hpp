// zmpm_simulation.hpp -- C++ host facade over the zmpm C ABI with the public
// members and virtual step structure of Ziran's MpmSimulationBase<T,3>
// (Lib/MPM/MpmSimulationBase.h:47-76 members, :150-231 virtuals) and the frame
// driver of SimulationBase (Lib/Ziran/Sim/SimulationBase.h:288-354), so scene
// code written against those names drives the gfx950 substep.
//
// Header-only; link with libzmpm.so.
#pragma once

#include "zmpm.h"
#include "zmpm_bgeo.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <functional>
#include <istream>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace zmpm {

// SimulationBase.h:21-29: thrown when a substep moved a particle more than dx
// and autorestart is on (MpmSimulationBase.cpp:1480-1499).
struct RestartException : public std::runtime_error {
    int frame;
    explicit RestartException(int f)
        : std::runtime_error("particle traveled more than a grid cell"), frame(f) {}
};

// The critical stress the reference derives from a stretch percentage when it
// attaches the phase field (addFBasedMpmForceWithAnisotropicPhaseField,
// MpmParticleHandleBase.cpp:175-222): the undegraded model's strain energy at
// F = (1 + p) I, where Q = I, R = (1 + p) I and J = (1 + p)^3.
//   QRStableNeoHookean::getSigmaCrit (QRStableNeoHookean.h:66-78), no fibres;
//   QRAnisotropic::getSigmaCrit (QRAnisotropic.h:65-77), fibres at mu;
//   NeoHookeanBorden::getSigmaCrit (NeoHookeanBorden.cpp:36-43), g = 1, no
//   deviatoric energy at a uniform stretch.
inline double sigmaCritFromPercentage(const zmpm_material& m, double percentage)
{
    const double s = 1.0 + percentage, J = s * s * s;
    const double vol = 0.5 * m.lambda * (J - 1.0) * (J - 1.0);
    switch (m.model) {
    case ZMPM_MODEL_QR_STABLE_NEO_HOOKEAN:
        return vol + m.mu / 2.0 * (3.0 * s * s - 3.0) - m.mu * (J - 1.0);
    case ZMPM_MODEL_QR_ANISOTROPIC:
        return vol + 1.5 * m.mu * (s - 1.0) * (s - 1.0);
    case ZMPM_MODEL_NEO_HOOKEAN_BORDEN: {
        const double kappa = 0.666666 * m.mu + m.lambda;
        return kappa / 2.0 * ((J * J - 1.0) / 2.0 - std::log(J));
    }
    default:
        throw std::invalid_argument("sigmaCritFromPercentage: only the phase-field models define getSigmaCrit");
    }
}

struct ZmpmError : public std::runtime_error {
    zmpm_status status;
    ZmpmError(zmpm_status s, const std::string& what)
        : std::runtime_error(what), status(s) {}
};

using TV = std::array<double, 3>;
using TM = std::array<double, 9>; // row-major

class MpmSimulationHip {
public:
    // ---- MpmSimulationBase public parameters
    double dx = 1;
    double dt = 1.0 / 24;
    double cfl = 0.6;
    double apic_rpic_ratio = 1;
    // MpmSimulationBase::TransferScheme (MpmSimulationBase.h:47-52)
    zmpm_transfer transfer_scheme = ZMPM_APIC_BLEND_RPIC;
    double flip_pic_ratio = 0;
    bool mls_mpm = false;
    int rpic_damping_iteration = 0; // MpmSimulationBase.h:135-136
    int pic_damping_iteration = 0;
    // AnisoFractureSimulation::implicit (AnisoFractureSimulation.h:61): phase
    // field by the CG solve of solveDamageSystem instead of the explicit update
    bool implicit = false;
    // FJMixedMpmForceHelper::state (FJMixedMpmForceHelper.h:38-40): 0 SOLVE_F,
    // 1 SOLVE_J; setFJState applies it to a running simulation
    int fj_state = 0;
    TV gravity { 0, 0, 0 };
    bool autorestart = true;
    // ---- device configuration (no counterpart in the reference)
    zmpm_precision precision = ZMPM_PREC_FP64;
    int domain_nodes = 256;
    int device = 0;
    // ---- particle attributes (Particles m/X/V, "C", "F", "element measure",
    //      AnisotropicPhaseField d, model g), host copies in upload order
    std::vector<TV> X, V;
    std::vector<double> mass, element_measure;
    std::vector<TM> C, F;
    std::vector<int32_t> material;
    std::vector<double> damage, degradation, laplacian;
    std::vector<double> plastic_state; // von Mises yield stress / Drucker-Prager log Jp / snow Jp per particle
    std::vector<double> max_dtilde;    // AnisotropicPhaseField::maxDTilde (implicit damage)
    std::vector<double> J;             // the "J" attribute of FJ-mixed particles (FJMixedMpmForceHelper.h:36)
    // ---- scene
    std::vector<zmpm_material> materials;
    std::vector<zmpm_collider> collision_objects;
    // ---- SimulationBase frame driver state
    int frame = 0;
    double time = 0;
    double frame_dt = 1.0 / 24;
    int end_frame = 0;
    // TimeStepping (Lib/Ziran/Sim/TimeStepping.h:15-33): dt bounds of the frame
    // driver; max_dt_original keeps the constructor's value (the default frame_dt)
    double min_dt = 1e-6;
    double max_dt = 1.0 / 24;
    double max_dt_original = 1.0 / 24;
    double time_since_last_frame = 0;
    int start_frame = 0;
    std::vector<std::function<void(int, int, double, double)>> begin_time_step_callbacks;
    // SimulationBase::restart_callbacks (SimulationBase.h:334-343): run once after a
    // restart has read its frame (the CFL autorestart shrinks max_dt through one)
    std::vector<std::function<void(int)>> restart_callbacks;
    int restarts = 0; // restarts taken by simulate() (no counterpart in the reference)

    MpmSimulationHip() = default;
    MpmSimulationHip(const MpmSimulationHip&) = delete;
    MpmSimulationHip& operator=(const MpmSimulationHip&) = delete;
    virtual ~MpmSimulationHip()
    {
        if (h_) zmpm_destroy(h_);
    }

    // MpmSimulationBase::addCollisionObject (MpmSimulationBase.cpp:87-91)
    void addCollisionObject(const zmpm_collider& c)
    {
        collision_objects.push_back(c);
        dirty_ = true;
    }

    void setFJState(int state)
    {
        fj_state = state;
        if (h_) check(zmpm_set_fj_state(h_, state), "zmpm_set_fj_state");
    }

    int addMaterial(const zmpm_material& m)
    {
        materials.push_back(m);
        dirty_ = true;
        return (int)materials.size() - 1;
    }

    // Particle with mass rho V / N and element measure V / N (MpmInitializationHelper.h:637-647).
    void addParticle(const TV& x, double m, double V0, int mat, const TM& F0 = { 1, 0, 0, 0, 1, 0, 0, 0, 1 })
    {
        X.push_back(x);
        V.push_back({ 0, 0, 0 });
        mass.push_back(m);
        element_measure.push_back(V0);
        C.push_back({ 0, 0, 0, 0, 0, 0, 0, 0, 0 });
        F.push_back(F0);
        material.push_back(mat);
        damage.push_back(0);
        degradation.push_back(1);
        laplacian.push_back(0);
        max_dtilde.push_back(0);
        J.push_back(1); // addFJMixedMpmForce starts J at 1
        const int pl = mat >= 0 && mat < (int)materials.size() ? materials[(size_t)mat].plasticity : ZMPM_PLASTICITY_NONE;
        plastic_state.push_back(pl == ZMPM_PLASTICITY_VON_MISES_STVK_HENCKY ? materials[(size_t)mat].yield_stress
                                    : pl == ZMPM_PLASTICITY_SNOW              ? 1.0 // SnowPlasticity::Jp starts at 1
                                                                              : 0.0);
        dirty_ = true;
    }

    // MpmSimulationBase::initialize (MpmSimulationBase.cpp:98-146): D^-1 = 4/dx^2
    // lives on the device; here the handle is created and the state uploaded.
    virtual void initialize()
    {
        if (h_) {
            zmpm_destroy(h_);
            h_ = nullptr;
        }
        zmpm_scene_desc d {};
        d.precision = precision;
        d.domain_nodes = domain_nodes;
        d.dx = dx;
        for (int a = 0; a < 3; a++) d.gravity[a] = gravity[a];
        d.apic_rpic_ratio = apic_rpic_ratio;
        d.cfl = cfl;
        d.device = device;
        d.transfer_scheme = transfer_scheme;
        d.flip_pic_ratio = flip_pic_ratio;
        d.mls_mpm = mls_mpm ? 1 : 0;
        d.rpic_damping_iteration = rpic_damping_iteration;
        d.pic_damping_iteration = pic_damping_iteration;
        d.implicit_damage = implicit ? 1 : 0;
        check(zmpm_create(&d, &h_), "zmpm_create");
        check(zmpm_set_materials(h_, (int32_t)materials.size(), materials.data()), "zmpm_set_materials");
        check(zmpm_set_fj_state(h_, fj_state), "zmpm_set_fj_state");
        for (const auto& c : collision_objects) check(zmpm_add_collider(h_, &c, nullptr), "zmpm_add_collider");
        upload();
    }

    // Binning, sort and grid activation run on the device at the start of every
    // substep (sortParticlesAndPolluteGrid, MpmSimulationBase.cpp:1531-1573).
    virtual void reinitialize()
    {
        if (dirty_) initialize();
    }

    // MpmSimulationBase::beginTimeStep (MpmSimulationBase.cpp:238-245)
    virtual void beginTimeStep(int frame_, int substep, double time_, double dt_)
    {
        dt = dt_;
        for (auto& f : begin_time_step_callbacks) f(frame_, substep, time_, dt_);
    }

    // MpmSimulationBase::advanceOneTimeStep (MpmSimulationBase.cpp:262-357):
    // one zmpm_substep on the device.  With autorestart the CFL condition is
    // checked after the substep (one device synchronisation); a particle that
    // moved more than dx throws RestartException for the last finished frame,
    // one frame earlier when max_dt cannot shrink any further, and otherwise
    // leaves a restart callback that halves max_dt (no lower than min_dt)
    // (MpmSimulationBase.cpp:1480-1499).  The device state of the violating
    // substep is the one a restart discards.
    virtual void advanceOneTimeStep(double dt_)
    {
        reinitialize();
        check(zmpm_substep(h_, dt_, time), "zmpm_substep");
        host_stale_ = true;
        if (autorestart) {
            zmpm_status s = zmpm_synchronize(h_);
            if (s == ZMPM_E_CFL) {
                // `frame` counts finished frames here: the reference's frame - 1
                int frame_to_restart = frame;
                if (max_dt <= min_dt) {
                    frame_to_restart--;
                    if (frame_to_restart < start_frame)
                        throw ZmpmError(ZMPM_E_CFL, "Unstable initial conditions detected, shrink min_dt or change other parameters");
                }
                else {
                    const double new_max_dt = std::max(max_dt / 2, min_dt);
                    restart_callbacks.push_back([new_max_dt, this](int) { max_dt = new_max_dt; });
                }
                throw RestartException(frame_to_restart);
            }
            check(s, "zmpm_synchronize");
        }
    }

    // `steps` substeps of size dt_ without host round trips.
    void advance(int steps, double dt_)
    {
        reinitialize();
        check(zmpm_run(h_, steps, dt_, time), "zmpm_run");
        time += steps * dt_;
        host_stale_ = true;
    }

    // MpmSimulationBase::calculateDt (MpmSimulationBase.cpp:1127-1151): cfl dx /
    // largest particle speed, max_dt when at rest (device reduction).
    virtual double calculateDt()
    {
        if (!h_) return max_dt;
        double d = max_dt;
        check(zmpm_calculate_dt(h_, max_dt, &d, nullptr, nullptr, nullptr), "zmpm_calculate_dt");
        return d;
    }

    // TimeStepping::nextDt (TimeStepping.h:45-59): max_dt no larger than
    // max_dt_original, clamp to [min_dt, max_dt] and land exactly on the frame
    // boundary without a tiny last substep.
    double nextDt(double d)
    {
        if (!(d > 0)) throw ZmpmError(ZMPM_E_ARG, "nextDt: dt must be positive");
        max_dt = std::min(max_dt, max_dt_original);
        d = (d < min_dt) ? min_dt : (d > max_dt) ? max_dt : d;
        if (time_since_last_frame + d >= frame_dt)
            d = frame_dt - time_since_last_frame;
        else if (time_since_last_frame + 2 * d > frame_dt)
            d = (frame_dt - time_since_last_frame) / 2;
        return d;
    }

    // SimulationBase::advanceOneFrame (SimulationBase.h:288-322): substeps of
    // nextDt(calculateDt()) until the frame time is reached (TimeStepping::advance).
    virtual void advanceOneFrame()
    {
        int substep = 0;
        for (;;) {
            const double h = nextDt(calculateDt());
            beginTimeStep(frame, substep, time, h);
            advanceOneTimeStep(h);
            time += h;
            time_since_last_frame += h;
            substep++;
            if (time_since_last_frame >= frame_dt) {
                time_since_last_frame = 0;
                break;
            }
        }
        substeps_last_frame = substep;
        frame++;
    }
    int substeps_last_frame = 0;

    // SimulationBase::simulate (SimulationBase.h:327-354).  With write_frames the
    // state file of the start frame and of every finished frame is written, as
    // the reference always does (write(outputFileName()), :332-333 and :319).
    // A RestartException from a frame rewinds to the frame it names (restart:
    // its state file, or with write_frames off the copy kept in memory of the
    // last two finished frames), runs the restart callbacks and carries on.
    bool write_frames = false;
    virtual void simulate()
    {
        frameFinished();
        while (frame < end_frame) {
            try {
                advanceOneFrame();
            }
            catch (const RestartException& e) {
                restart(e.frame);
                for (auto& f : restart_callbacks) f(frame);
                restart_callbacks.clear();
                restarts++;
                continue;
            }
            frameFinished();
        }
        sync();
    }

    // Copies the device state back into the host attribute arrays.
    void sync()
    {
        if (!h_ || !host_stale_) return;
        zmpm_particles p = view();
        zmpm_status s = zmpm_download(h_, &p);
        if (s != ZMPM_OK && s != ZMPM_E_CFL) check(s, "zmpm_download");
        host_stale_ = false;
    }

    // Restart state: the particle attributes in upload order plus the clock,
    // as one binary record (the reference streams its DataManager attributes and
    // scene, MpmSimulationBase.cpp:1060-1114; this is the same role in a flat
    // layout of its own: magic, version, n, frame, time, time_since_last_frame,
    // then each attribute array).  Materials and colliders are scene setup and
    // are not part of it, as in the reference's restart files.
    static constexpr char kStateMagic[8] = { 'Z', 'M', 'P', 'M', 'S', 'T', 'A', 'T' };
    static constexpr int32_t kStateVersion = 2;

    virtual void writeState(std::ostream& out)
    {
        sync();
        const int64_t n = (int64_t)X.size();
        out.write(kStateMagic, sizeof(kStateMagic));
        out.write((const char*)&kStateVersion, sizeof(kStateVersion));
        out.write((const char*)&n, sizeof(n));
        out.write((const char*)&frame, sizeof(frame));
        out.write((const char*)&time, sizeof(time));
        out.write((const char*)&time_since_last_frame, sizeof(time_since_last_frame));
        forEachAttribute([&](char* p, size_t bytes) { out.write(p, (std::streamsize)bytes); });
        if (!out) throw ZmpmError(ZMPM_E_ARG, "writeState: write failed");
    }

    // SimulationBase::readState counterpart: replaces the particle state and the
    // clock; the device copy is refreshed before the next substep.
    virtual void readState(std::istream& in)
    {
        char magic[sizeof(kStateMagic)];
        int32_t version = 0;
        int64_t n = -1;
        in.read(magic, sizeof(magic));
        in.read((char*)&version, sizeof(version));
        in.read((char*)&n, sizeof(n));
        if (!in || std::memcmp(magic, kStateMagic, sizeof(magic)) != 0 || version != kStateVersion || n < 0)
            throw ZmpmError(ZMPM_E_ARG, "readState: not a state record of this version");
        in.read((char*)&frame, sizeof(frame));
        in.read((char*)&time, sizeof(time));
        in.read((char*)&time_since_last_frame, sizeof(time_since_last_frame));
        resizeAttributes((size_t)n);
        forEachAttribute([&](char* p, size_t bytes) { in.read(p, (std::streamsize)bytes); });
        if (!in) throw ZmpmError(ZMPM_E_ARG, "readState: truncated state record");
        for (int32_t m : material)
            if (m < 0 || m >= (int32_t)materials.size()) throw ZmpmError(ZMPM_E_ARG, "readState: material index out of range");
        host_stale_ = false;
        dirty_ = true;
    }

    // SimulationBase::write / read / outputFileName (SimulationBase.h:148-200)
    // and restart (:265-285): state files output_dir/restart_<frame>.dat.
    std::string output_dir = ".";
    virtual std::string outputFileName(const std::string& name = "restart", const std::string& suffix = ".dat") const
    {
        return name + "_" + std::to_string(frame) + suffix;
    }
    virtual void write(const std::string& filename)
    {
        std::ofstream f(output_dir + "/" + filename, std::ios::binary);
        if (!f) throw ZmpmError(ZMPM_E_ARG, "write: cannot open " + output_dir + "/" + filename);
        writeState(f);
        if (write_partio) writePartio(output_dir + "/" + outputFileName("partio", ".bgeo"));
    }

    // MpmSimulationBase::write_partio and writePartio (PartioIO.h:166-200): the
    // particle positions as a BGEO point file, output_dir/partio_<frame>.bgeo,
    // next to each state file.
    bool write_partio = false;
    virtual void writePartio(const std::string& path)
    {
        sync();
        try {
            writeBgeoPoints(path, (int64_t)X.size(), X.empty() ? nullptr : X[0].data(), {});
        }
        catch (const std::exception& e) {
            throw ZmpmError(ZMPM_E_ARG, e.what());
        }
    }

    // The AnisoMPM visualizer's particle file (aniso_visualize_particles_vec,
    // AnisoVisualizer.h:38-63, without its fibre attribute): positions, damage
    // "d" and "laplacian".
    virtual void writePartioDamage(const std::string& path)
    {
        sync();
        const int64_t n = (int64_t)X.size();
        if ((int64_t)damage.size() != n || (int64_t)laplacian.size() != n)
            throw ZmpmError(ZMPM_E_ARG, "writePartioDamage: no phase field");
        try {
            writeBgeoPoints(path, n, n ? X[0].data() : nullptr,
                            { BgeoAttribute { "d", 1, damage.data() }, BgeoAttribute { "laplacian", 1, laplacian.data() } });
        }
        catch (const std::exception& e) {
            throw ZmpmError(ZMPM_E_ARG, e.what());
        }
    }
    virtual void read(const std::string& filename)
    {
        std::ifstream f(output_dir + "/" + filename, std::ios::binary);
        if (!f) throw ZmpmError(ZMPM_E_ARG, "read: cannot open " + output_dir + "/" + filename);
        readState(f);
    }
    // SimulationBase::restart (SimulationBase.h:265-285) with the read simulate()
    // does next: the state of frame new_restart_frame (its state file, else the
    // in-memory copy simulate() keeps) and the clock of that frame.
    virtual void restart(int new_restart_frame)
    {
        frame = new_restart_frame;
        time = frame * frame_dt; // TimeStepping::reset (TimeStepping.h:88-92)
        time_since_last_frame = 0;
        bool from_memory = false;
        for (const auto& kv : kept_frames_)
            if (kv.first == new_restart_frame && !write_frames) {
                std::istringstream in(kv.second);
                readState(in);
                from_memory = true;
            }
        if (!from_memory) read(outputFileName());
        if (frame != new_restart_frame) throw ZmpmError(ZMPM_E_ARG, "restart: file holds another frame");
    }

    zmpm_handle handle() const { return h_; }

private:
    zmpm_handle h_ = nullptr;
    bool dirty_ = true;
    bool host_stale_ = false;
    // state records of the last two finished frames, for autorestart without state files
    std::vector<std::pair<int, std::string>> kept_frames_;

    // End of a frame in simulate(): its state file (write_frames) and, for
    // autorestart without files, an in-memory copy.
    void frameFinished()
    {
        if (write_frames) write(outputFileName());
        if (autorestart && !write_frames) {
            std::ostringstream out;
            writeState(out);
            if (kept_frames_.size() == 2) kept_frames_.erase(kept_frames_.begin());
            kept_frames_.emplace_back(frame, out.str());
        }
    }

    template <class Fn>
    void forEachAttribute(Fn&& fn)
    {
        const size_t n = X.size();
        fn((char*)X.data(), n * sizeof(TV));
        fn((char*)V.data(), n * sizeof(TV));
        fn((char*)C.data(), n * sizeof(TM));
        fn((char*)F.data(), n * sizeof(TM));
        fn((char*)mass.data(), n * sizeof(double));
        fn((char*)element_measure.data(), n * sizeof(double));
        fn((char*)material.data(), n * sizeof(int32_t));
        fn((char*)damage.data(), n * sizeof(double));
        fn((char*)degradation.data(), n * sizeof(double));
        fn((char*)laplacian.data(), n * sizeof(double));
        fn((char*)plastic_state.data(), n * sizeof(double));
        fn((char*)max_dtilde.data(), n * sizeof(double));
        fn((char*)J.data(), n * sizeof(double));
    }

    void resizeAttributes(size_t n)
    {
        X.resize(n);
        V.resize(n);
        C.resize(n);
        F.resize(n);
        mass.resize(n);
        element_measure.resize(n);
        material.resize(n);
        damage.resize(n);
        degradation.resize(n);
        laplacian.resize(n);
        plastic_state.resize(n);
        max_dtilde.resize(n);
        J.resize(n);
    }

    zmpm_particles view()
    {
        zmpm_particles p {};
        p.n = (int64_t)X.size();
        p.x = X.empty() ? nullptr : X[0].data();
        p.v = V.empty() ? nullptr : V[0].data();
        p.C = C.empty() ? nullptr : C[0].data();
        p.F = F.empty() ? nullptr : F[0].data();
        p.m = mass.data();
        p.V0 = element_measure.data();
        p.material = material.data();
        p.d = damage.data();
        p.g = degradation.data();
        p.laplacian = laplacian.data();
        p.plastic_state = plastic_state.data();
        p.max_dtilde = max_dtilde.data();
        p.J = J.data();
        return p;
    }

    void upload()
    {
        zmpm_particles p = view();
        check(zmpm_upload(h_, &p), "zmpm_upload");
        dirty_ = false;
        host_stale_ = false;
    }

    void check(zmpm_status s, const char* where)
    {
        if (s == ZMPM_OK) return;
        std::string msg = std::string(where) + ": status " + std::to_string((int)s);
        if (h_) msg += std::string(" (") + zmpm_last_error(h_) + ")";
        throw ZmpmError(s, msg);
    }
};

} // namespace zmpm
