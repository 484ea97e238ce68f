// trajectory.hpp -- binary trajectory frames in the reference's format (host C++, no GPU).
//
// Frame layout (writer: serial/box.h:549-568 and parallel/box.h:476-502; reader:
// analyze/read.h:215-274), little-endian, no padding:
//   f64 t, f64 Lx, f64 Ly, i32 nb, i32 Ns, i32 Nm,
//   Nm x {f32 x, f32 y}            membrane beads, unwrapped position (r_original)
//   Ns x {f32 x, f32 y, f32 theta}  swimmers / particles
// RepulsiveParticle and ActiveBrownianChain write r_original (particle.h:759-764, 1038-1045);
// BasicDynamicParticle (the continuous model) writes the wrapped r (particle.h:123-128).
// An ActiveBrownianChain with F0 == 0 writes no theta (particle.h:1038-1039); frames are then
// 36 + 8 Nm + 8 Ns bytes instead of 36 + 8 Nm + 12 Ns.
#pragma once

#include <cstdint>
#include <cstring>
#include <istream>
#include <ostream>
#include <string>
#include <vector>

namespace spp_traj {

struct Frame {
    double t = 0, Lx = 0, Ly = 0;
    int32_t nb = 1;
    std::vector<float> mx, my;          // Nm membrane beads
    std::vector<float> x, y, theta;     // Ns particles (theta empty when not written)
    bool healthy = false;
    int32_t Ns() const { return (int32_t)x.size(); }
    int32_t Nm() const { return (int32_t)mx.size(); }
};

inline void put(std::ostream& os, const void* p, size_t n) { os.write((const char*)p, (std::streamsize)n); }

inline void write_frame(std::ostream& os, const Frame& f)
{
    int32_t Ns = f.Ns(), Nm = f.Nm();
    put(os, &f.t, 8);
    put(os, &f.Lx, 8);
    put(os, &f.Ly, 8);
    put(os, &f.nb, 4);
    put(os, &Ns, 4);
    put(os, &Nm, 4);
    for (int32_t i = 0; i < Nm; ++i) {
        put(os, &f.mx[i], 4);
        put(os, &f.my[i], 4);
    }
    bool th = !f.theta.empty();
    for (int32_t i = 0; i < Ns; ++i) {
        put(os, &f.x[i], 4);
        put(os, &f.y[i], 4);
        if (th) put(os, &f.theta[i], 4);
    }
}

// Reads one frame.  Health follows Scene's reader (read.h:230-271): a truncated frame, t outside
// [0, 1e6] or a particle count outside [0, 1e6] marks it unhealthy.  with_theta tells whether
// the particles carry a heading (the file does not say).
inline bool read_frame(std::istream& is, Frame& f, bool with_theta = true)
{
    f.healthy = false;
    int32_t Ns = 0, Nm = 0;
    is.read((char*)&f.t, 8);
    is.read((char*)&f.Lx, 8);
    is.read((char*)&f.Ly, 8);
    is.read((char*)&f.nb, 4);
    is.read((char*)&Ns, 4);
    is.read((char*)&Nm, 4);
    if (!is || Ns < 0 || Ns > 1000000 || Nm < 0 || Nm > 1000000) return false;
    f.mx.assign(Nm, 0.f);
    f.my.assign(Nm, 0.f);
    f.x.assign(Ns, 0.f);
    f.y.assign(Ns, 0.f);
    f.theta.assign(with_theta ? Ns : 0, 0.f);
    for (int32_t i = 0; i < Nm; ++i) {
        is.read((char*)&f.mx[i], 4);
        is.read((char*)&f.my[i], 4);
    }
    for (int32_t i = 0; i < Ns; ++i) {
        is.read((char*)&f.x[i], 4);
        is.read((char*)&f.y[i], 4);
        if (with_theta) is.read((char*)&f.theta[i], 4);
    }
    if (!is) return false;
    f.healthy = !(f.t < 0 || f.t > 1000000);
    return true;
}

// fix-file semantics (analyze/fix-file.cpp:11-36): every complete, healthy frame in order; a
// truncated tail (a run killed mid-write) is dropped.
inline std::vector<Frame> read_healthy_frames(std::istream& is, bool with_theta = true)
{
    std::vector<Frame> out;
    Frame f;
    while (is.peek() != EOF) {
        if (!read_frame(is, f, with_theta)) break;
        if (f.healthy) out.push_back(f);
    }
    return out;
}

}  // namespace spp_traj
