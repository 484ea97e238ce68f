// box.hpp -- C++ host facade with the reference's Box call surface over the engine's C ABI.
//
// The reference drives its particle path through Box methods (serial/box.h:40-57,
// parallel/box.h:40-53): Interact, Move / Move_Runge_Kutta{2,4}_k, One_Step, Multi_Step,
// Update_Cells, Save / Load, Save_Polarization and the binary frame operators.  This class keeps
// those names and meanings; every call forwards to include/spp_engine.h, which runs the HIP
// kernels.  Errors throw spp_gpu::Error (the reference prints and calls exit(0)).
#pragma once

#include <algorithm>
#include <cmath>
#include <fstream>
#include <istream>
#include <ostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/spp_engine.h"
#include "membrane.hpp"
#include "trajectory.hpp"

namespace spp_gpu {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// State_Hyper_Vector (shared/state-hyper-vector.h:6-258) on the host: a box's state as the 3N
// vector (r, theta) with the reference's arithmetic -- +, -, scaling and their compound forms end
// in Periodic_Transform (minimum-image positions, theta into (-pi, pi], :83-150, 218-225); the
// product is sum(r.r' + theta theta') (:157-167); Square wraps theta first (:201-211).  It also
// carries
//   * the unwrapped positions, which follow the same arithmetic without the wrap, so a deviation
//     added to a state moves r_original with r (the reference's vector holds r only, and Load leaves
//     r_original as it was, parallel/box.h:229-235, which loses a positional deviation after a step);
//   * the clock (spp_clock): the noise stream's step, the counter-based stand-in for the generator
//     copy the reference keeps with each vector (:43-57, 169-177), the time, and the displacement
//     policy's rebuild estimates, so that a Load replays the saved run bit for bit.
class State {
public:
    int N = 0;
    double Lx = 0, Ly = 0;                   // half widths of the box the positions wrap in
    std::vector<double> x, y, theta, xu, yu;
    spp_clock clock{};  // noise step, time and list-rebuild schedule (exact replay)

    State() = default;
    State(int n, double lx, double ly) : N(n), Lx(lx), Ly(ly), x(n), y(n), theta(n), xu(n), yu(n) {}

    State operator+(const State& s) const { return combine(s, +1.0); }
    State operator-(const State& s) const { return combine(s, -1.0); }
    State& operator+=(const State& s) { return *this = combine(s, +1.0); }
    State& operator-=(const State& s) { return *this = combine(s, -1.0); }
    State operator*(double f) const { return scaled(f, false); }
    State operator/(double f) const { return scaled(f, true); }
    double operator*(const State& s) const
    {
        check_size(s);
        double r = 0;
        for (int i = 0; i < N; ++i) {
            r += x[i] * s.x[i] + y[i] * s.y[i];
            r += theta[i] * s.theta[i];
        }
        return r;
    }
    void Null()
    {
        for (auto* v : {&x, &y, &theta, &xu, &yu}) std::fill(v->begin(), v->end(), 0.0);
    }
    double Square() const
    {
        double r = 0;
        for (int i = 0; i < N; ++i) {
            r += x[i] * x[i] + y[i] * y[i];
            const double d = wrap_theta(theta[i]);
            r += d * d;
        }
        return r;
    }
    double Magnitude() const { return std::sqrt(Square()); }
    void Periodic_Transform()
    {
        for (int i = 0; i < N; ++i) {
            x[i] = wrap(x[i], 2 * Lx);
            y[i] = wrap(y[i], 2 * Ly);
            theta[i] = wrap_theta(theta[i]);
        }
    }
    void Unit() { *this = *this / Magnitude(); }
    int Max_Index() const  // largest wrapped |theta| (state-hyper-vector.h:228-245)
    {
        int index = 0;
        double mx = 0;
        for (int i = 0; i < N; ++i) {
            const double d = std::fabs(wrap_theta(theta[i]));
            if (d > mx) {
                mx = d;
                index = i;
            }
        }
        return index;
    }

private:
    // C2DVector::Periodic_Transform, one component (c2dvector.h:67-73)
    static double wrap(double v, double L2) { return v - L2 * (double)(int)std::floor(v / L2 + 0.5); }
    static double wrap_theta(double th) { return th - 2 * M_PI * std::ceil((th - M_PI) / (2 * M_PI)); }
    void check_size(const State& s) const
    {
        if (s.N != N) throw Error(SPP_ERR_ARG, "state vectors hold different particle numbers");
    }
    State combine(const State& s, double sign) const  // the result keeps this vector's clock
    {
        check_size(s);
        State r = *this;
        for (int i = 0; i < N; ++i) {
            r.x[i] = x[i] + sign * s.x[i];
            r.y[i] = y[i] + sign * s.y[i];
            r.theta[i] = theta[i] + sign * s.theta[i];
            r.xu[i] = xu[i] + sign * s.xu[i];
            r.yu[i] = yu[i] + sign * s.yu[i];
        }
        r.Periodic_Transform();
        return r;
    }
    State scaled(double f, bool divide) const
    {
        State r = *this;
        for (int i = 0; i < N; ++i) {
            r.x[i] = divide ? x[i] / f : x[i] * f;
            r.y[i] = divide ? y[i] / f : y[i] * f;
            r.theta[i] = divide ? theta[i] / f : theta[i] * f;
            r.xu[i] = divide ? xu[i] / f : xu[i] * f;
            r.yu[i] = divide ? yu[i] / f : yu[i] * f;
        }
        r.Periodic_Transform();
        return r;
    }
};

class Box {
public:
    double t = 0;             // simulated time (reference Box::t)
    int N = 0, Ns = 0, Nm = 0;
    double polarization = 0;  // last Compute_Polarization

    explicit Box(const spp_params& p) : params_(p)
    {
        check(spp_create(&params_, &e_), nullptr);
        N = p.n;
        Nm = p.n_membrane;
        Ns = N - Nm;
    }
    ~Box() { spp_destroy(e_); }
    Box(const Box&) = delete;
    Box& operator=(const Box&) = delete;

    spp_engine* engine() { return e_; }
    const spp_params& params() const { return params_; }

    // Box::Init with given positions and headings (the reference places them with
    // Random_Formation / Triangle_Lattice_Formation, shared/set-up.h); unwrapped = wrapped.
    void Init(const std::vector<double>& x, const std::vector<double>& y, const std::vector<double>& theta)
    {
        if ((int)x.size() != N || (int)y.size() != N || (int)theta.size() != N)
            throw Error(SPP_ERR_ARG, "Box::Init needs one x, y and theta per particle");
        check(spp_set_state(e_, x.data(), y.data(), nullptr, nullptr, theta.data(), N));
        set_time_origin(0.0);
    }

    // Box::Positioning_Particles (serial/box.h:211-272): resume from the last healthy frame of a
    // trajectory file (truncated tails and unhealthy frames are skipped, as fix-file does).
    // Returns false if the file holds no usable frame.
    bool Positioning_Particles(const std::string& name)
    {
        std::ifstream is(name, std::ios::binary);
        if (!is) return false;
        auto frames = spp_traj::read_healthy_frames(is, writes_theta());
        if (frames.empty() || frames.back().Ns() + frames.back().Nm() != N) return false;
        load_frame(frames.back());
        return true;
    }

    // operator>>(istream&, Box*), parallel/box.h:505-524: reads the next frame and loads its time,
    // positions and headings.  A frame that cannot be read whole, or holds another particle count,
    // sets the stream's failbit and leaves the box as it was.
    friend std::istream& operator>>(std::istream& is, Box* box)
    {
        spp_traj::Frame f;
        if (!spp_traj::read_frame(is, f, box->writes_theta()) || f.Ns() + f.Nm() != box->N) {
            is.setstate(std::ios::failbit);
            return is;
        }
        box->load_frame(f);
        return is;
    }

    void Update_Cells() { check(spp_build_neighbors(e_)); }        // serial/box.h:107-134
    void Update_Neighbor_List() { check(spp_build_neighbors(e_)); } // serial/box.h:137-179
    void Interact() { check(spp_interact(e_)); }                    // serial/box.h:322-371
    void Move() { stage(0); }                                       // :374-378
    void Move_Runge_Kutta2_1() { stage(1); }                        // :382-386
    void Move_Runge_Kutta2_2() { stage(2); }                        // :389-393
    void Move_Runge_Kutta4_1() { stage(1); }                        // :398-402
    void Move_Runge_Kutta4_2() { stage(2); }
    void Move_Runge_Kutta4_3() { stage(3); }
    void Move_Runge_Kutta4_4() { stage(4); }                        // :419-423
    void One_Step() { check(spp_one_step(e_)); sync_time(); }       // :427-455
    void Multi_Step(int steps) { check(spp_multi_step(e_, steps)); sync_time(); }  // :469-506
    // :509-515 runs steps/interval chunks then the remainder
    void Multi_Step(int steps, int interval)
    {
        for (int i = 0; interval > 0 && i < steps / interval; ++i) Multi_Step(interval);
        Multi_Step(interval > 0 ? steps % interval : steps);
    }

    void Save(int slot) { check(spp_snapshot(e_, slot)); }          // parallel/box.h:246-267
    void Load(int slot) { check(spp_restore(e_, slot)); sync_time(); }  // :222-243

    // Box::Save(State_Hyper_Vector&) (parallel/box.h:246-267): the state and the noise clock into a
    // host vector; re-bins first like the reference, so a Load of it reproduces the same order.
    void Save(State& sv)
    {
        if (sv.N != N) throw Error(SPP_ERR_ARG, "Box::Save: the state vector holds another particle number");
        check(spp_build_neighbors(e_));
        Get_State(sv.x, sv.y, sv.xu, sv.yu, sv.theta);
        check(spp_get_clock(e_, &sv.clock));
        sv.clock.t = t;
        sv.Lx = params_.Lx;
        sv.Ly = params_.Ly;
    }
    // Box::Load(const State_Hyper_Vector&) (:222-243): positions, headings and the noise clock.
    void Load(const State& sv)
    {
        if (sv.N != N) throw Error(SPP_ERR_ARG, "Box::Load: the state vector holds another particle number");
        check(spp_set_state(e_, sv.x.data(), sv.y.data(), sv.xu.data(), sv.yu.data(), sv.theta.data(), N));
        check(spp_set_clock(e_, &sv.clock));
        set_time_origin(sv.clock.t);
    }

    // The Particle-level surface (shared/particle.h): particle(i).Interact(particle(j)) adds the
    // pair's force and torque to both (Particle::Interact, :728-757, 296-338, 964-1012);
    // Reset_Forces is Particle::Reset for all.  Moves stay Box-level (Move*, above), as every
    // reference caller applies them (cell.h:137-147).
    class Particle {
    public:
        void Interact(const Particle& other) const { box_->check(spp_pair_interact(box_->e_, i_, other.i_)); }
        int index() const { return i_; }

    private:
        friend class Box;
        Particle(Box* b, int i) : box_(b), i_(i) {}
        Box* box_;
        int i_;
    };
    Particle particle(int i)
    {
        if (i < 0 || i >= N) throw Error(SPP_ERR_ARG, "particle index out of range");
        return Particle(this, i);
    }
    void Reset_Forces() { check(spp_reset_forces(e_)); }
    // every particle's accumulated force and torque (the reference's Particle::f and torque)
    void Get_Forces(std::vector<double>& fx, std::vector<double>& fy, std::vector<double>& torque)
    {
        fx.resize(N);
        fy.resize(N);
        torque.resize(N);
        check(spp_get_forces(e_, fx.data(), fy.data(), torque.data(), N));
    }

    void Compute_Polarization()                                     // serial/box.h:531-538
    {
        check(spp_polarization(e_, &polarization, nullptr, nullptr));
    }
    void Save_Polarization(std::ostream& os)                        // :541-546
    {
        Compute_Polarization();
        os << t << "\t" << polarization << std::endl;
    }

    // Vesicle observables (serial/beadbox.h:419-506): Save_Particles_Positions keeps the
    // unwrapped positions and headings as the reference for the next finite-difference velocity;
    // Compute_All_Variables evaluates the membrane/swimmer moments and saves the positions again;
    // Save_All_Variables appends one row (and the header, once, to an empty stream).
    MembraneVariables variables;

    void Save_Particles_Positions()
    {
        std::vector<double> xw, yw, th;
        Get_State(xw, yw, x_old_, y_old_, th);
        t_old_ = t;
    }
    void Compute_All_Variables()
    {
        if (Nm <= 0) throw Error(SPP_ERR_ARG, "Compute_All_Variables needs ring beads (n_membrane > 0)");
        if ((int)x_old_.size() != N) Save_Particles_Positions();
        std::vector<double> xw, yw, xu, yu, th;
        Get_State(xw, yw, xu, yu, th);
        variables = membrane_variables(xu.data(), yu.data(), x_old_.data(), y_old_.data(), Nm, Ns, t - t_old_);
        x_old_.swap(xu);
        y_old_.swap(yu);
        t_old_ = t;
    }
    void Save_All_Variables(std::ostream& os)
    {
        Compute_All_Variables();
        if (!header_written_) {
            // the header goes only into an empty file: an appending stream reports position 0
            // until its first write, so ask the buffer for the end (no stream state on failure)
            auto end = os.rdbuf() ? os.rdbuf()->pubseekoff(0, std::ios::end, std::ios::out) : std::streampos(-1);
            if (end == std::streampos(0)) write_variables_header(os);
            header_written_ = true;
        }
        write_variables_row(os, t, variables);
    }

    // Current state in caller order.
    void Get_State(std::vector<double>& xw, std::vector<double>& yw, std::vector<double>& xu,
                   std::vector<double>& yu, std::vector<double>& th)
    {
        xw.resize(N);
        yw.resize(N);
        xu.resize(N);
        yu.resize(N);
        th.resize(N);
        check(spp_get_state(e_, xw.data(), yw.data(), xu.data(), yu.data(), th.data(), N));
    }

    // operator<<(ostream&, Box*), serial/box.h:549-568: one binary frame.
    friend std::ostream& operator<<(std::ostream& os, Box* box)
    {
        std::vector<double> xw, yw, xu, yu, th;
        box->Get_State(xw, yw, xu, yu, th);
        spp_traj::Frame f;
        f.t = box->t;
        f.Lx = box->params_.Lx;
        f.Ly = box->params_.Ly;
        // the first swimmer's bead count (parallel/beadbox.h:463)
        f.nb = box->params_.model == SPP_MODEL_ABP && box->params_.nb_swimmer >= 2 ? box->params_.nb_swimmer : 1;
        // continuous particles write the wrapped r (BasicDynamicParticle::Write), the others
        // r_original (RepulsiveParticle::Write, ActiveBrownianChain::Write); swimmers without
        // self-propulsion write no heading (particle.h:1038-1045)
        bool wrapped = box->params_.model == SPP_MODEL_CONTINUOUS;
        bool with_theta = box->writes_theta();
        for (int i = 0; i < box->Nm; ++i) {
            f.mx.push_back((float)xu[i]);
            f.my.push_back((float)yu[i]);
        }
        for (int i = box->Nm; i < box->N; ++i) {
            f.x.push_back((float)(wrapped ? xw[i] : xu[i]));
            f.y.push_back((float)(wrapped ? yw[i] : yu[i]));
            if (with_theta) f.theta.push_back((float)th[i]);
        }
        spp_traj::write_frame(os, f);
        return os;
    }

    // Whether frames carry the particles' headings: all models except swimmers with F0 == 0.
    bool writes_theta() const { return !(params_.model == SPP_MODEL_ABP && params_.v0 == 0.0); }

private:
    void stage(int s) { check(spp_move_stage(e_, s)); sync_time(); }
    void set_time_origin(double t0)
    {
        spp_info inf;
        check(spp_get_info(e_, &inf));
        t0_ = t0;
        base_ = inf.t;
        t = t0;
    }
    void sync_time()
    {
        spp_info inf;
        if (spp_get_info(e_, &inf) == SPP_OK) t = t0_ + (inf.t - base_);
    }
    // A frame's positions (r_original for ring beads and, except for continuous particles, for
    // swimmers; r for continuous particles) become both positions, wrapped as the reference's
    // r = r_original; r.Periodic_Transform() (serial/box.h:245-258); headings default to 0.
    void load_frame(const spp_traj::Frame& f)
    {
        std::vector<double> xu(N), yu(N), xw(N), yw(N), th(N, 0.0);
        for (int i = 0; i < f.Nm(); ++i) {
            xu[i] = f.mx[i];
            yu[i] = f.my[i];
        }
        for (int i = 0; i < f.Ns(); ++i) {
            xu[f.Nm() + i] = f.x[i];
            yu[f.Nm() + i] = f.y[i];
            th[f.Nm() + i] = f.theta.empty() ? 0.0 : f.theta[i];
        }
        for (int i = 0; i < N; ++i) {
            xw[i] = xu[i] - 2 * params_.Lx * (double)(int)std::floor(xu[i] / (2 * params_.Lx) + 0.5);
            yw[i] = yu[i] - 2 * params_.Ly * (double)(int)std::floor(yu[i] / (2 * params_.Ly) + 0.5);
        }
        check(spp_set_state(e_, xw.data(), yw.data(), xu.data(), yu.data(), th.data(), N));
        set_time_origin(f.t);
    }

    void check(int rc, spp_engine* e) const
    {
        if (rc != SPP_OK) throw Error(rc, spp_last_error(e));
    }
    void check(int rc) const { check(rc, e_); }

    spp_params params_;
    spp_engine* e_ = nullptr;
    double t0_ = 0, base_ = 0;
    std::vector<double> x_old_, y_old_;
    double t_old_ = 0;
    bool header_written_ = false;
};

}  // namespace spp_gpu
