// TEST INFRASTRUCTURE ONLY (oracle). Never linked into the product, never run by it.
//
// Host driver for the reference's OWN update functions, used to write golden vectors. It includes
// the reference headers where they lie (include path -I<reference root>) and calls, per cell, the
// reference's device functions compiled for the host:
//   RCS                                 lbm.cuh:20-262
//   inlet_bc_calculation                jet/boundary_conditions.cuh:16-39
//   outlet_neumann_bc_calculation       jet/boundary_conditions.cuh:45-109
//   init_jet_bulk / init_jet_inlet      jet/jetIn.cuh:16-60
//   init_static_bubble                  bubble/bubbleIn.cuh:36-52
//   init_rti_domain                     rti/rtiIn.cuh:47-71
//   write_vti                           io/vtiWriter.cuh:92-145 (REF_VTI snapshots only)
// in the step order of main.cu:38-55 (RCS over interior cells, then the jet boundary kernel
// with (B, Aold), then the pointer swap), with the interior guards of kernels.cu:22-131.
//
// Build variants (oracle/Makefile):
//   default                  constants.cuh exactly as shipped (128^3 incl. ghosts, jet case).
//   -DREF_OVERRIDE_CONSTANTS  ref_constants.hpp supplies the constexpr block with -D tunables.
//   -DREF_PERIODIC_Y          y made periodic by routing pully/force_y_outlet_limited through the
//                            reference's own wrapy (bounds.cuh:49-62), which makes the y-ghost
//                            branch (lbm.cuh:79-151) unreachable. This is the only behavioural
//                            substitution and is used only for the periodic-bubble vectors.
//
// Output: raw float32, 11 fields in the order rhor rhob ux uy uz Pixx Pixy Piyy Piyz Pizz Pixz,
// each over the compact region x in [1,NX-2], z in [1,NZ-2] and y in [0,NY-1] (jet, ghost planes
// included) or y in [1,NY-2] (periodic y), x fastest.
#ifdef REF_OVERRIDE_CONSTANTS
#include "ref_constants.hpp"
#endif

// memory.cuh first: it brings <cuda_runtime.h> (memory.cuh:4), which defines the CUDA qualifiers
// the other reference headers use.
#include "memory.cuh"
#include "utilities/bounds.cuh"
#include "utilities/mathUtilities.cuh"

#ifdef REF_PERIODIC_Y
template <label_t I>
inline label_t ref_pully_periodic(const label_t y) noexcept { return wrapy(pully<I>(y)); }
inline label_t ref_force_y_periodic(const int y) noexcept { return wrapy(static_cast<label_t>(y)); }
#define pully ref_pully_periodic
#define force_y_outlet_limited ref_force_y_periodic
#endif

#include "lbm.cuh"
#include "jet/jetIn.cuh"
#include "jet/boundary_conditions.cuh"
#include "bubble/bubbleIn.cuh"
#include "rti/rtiIn.cuh"
#include "io/vtiWriter.cuh"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct HostState {
    std::vector<float> buf[11];
    MomentsDevice view{};
    HostState() {
        for (auto &b : buf) b.assign(Ncells, 0.0f);  // zero_device (memory.cuh:153-164)
        view.rhor = buf[0].data(); view.rhob = buf[1].data();
        view.ux = buf[2].data(); view.uy = buf[3].data(); view.uz = buf[4].data();
        view.Pixx = buf[5].data(); view.Pixy = buf[6].data(); view.Piyy = buf[7].data();
        view.Piyz = buf[8].data(); view.Pizz = buf[9].data(); view.Pixz = buf[10].data();
    }
};

bool g_periodic_y =
#ifdef REF_PERIODIC_Y
    true;
#else
    false;
#endif

label_t y_lo() { return g_periodic_y ? 1u : 0u; }
label_t y_hi() { return g_periodic_y ? NY - 2u : NY - 1u; }

// With REF_VTI set in the environment, every snapshot is also written by the reference's own VTI
// writer (io/vtiWriter.cuh:92-145, full extent incl. ghosts) to <path>.vti, and the five fields it
// prints are dumped over the full extent to <path>.full5.
void dump_vti(const HostState &s, const std::string &path) {
    LbmHost h{};
    h.rhor = const_cast<float *>(s.buf[0].data());
    h.rhob = const_cast<float *>(s.buf[1].data());
    h.ux = const_cast<float *>(s.buf[2].data());
    h.uy = const_cast<float *>(s.buf[3].data());
    h.uz = const_cast<float *>(s.buf[4].data());
    write_vti(std::filesystem::path(path + ".vti"), h);
    FILE *f = std::fopen((path + ".full5").c_str(), "wb");
    if (!f) { std::perror(path.c_str()); std::exit(2); }
    for (int fi = 0; fi < 5; ++fi) std::fwrite(s.buf[fi].data(), sizeof(float), Ncells, f);
    std::fclose(f);
}

void dump(const HostState &s, const std::string &path) {
    if (std::getenv("REF_VTI")) dump_vti(s, path);
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror(path.c_str()); std::exit(2); }
    std::vector<float> row;
    for (int fi = 0; fi < 11; ++fi)
        for (label_t z = 1; z + 1 < NZ; ++z)
            for (label_t y = y_lo(); y <= y_hi(); ++y) {
                const float *p = s.buf[fi].data() + idx(1, y, z);
                std::fwrite(p, sizeof(float), NX - 2, f);
            }
    std::fclose(f);
}

// splitmix64 -> [0,1)
struct Rng {
    uint64_t s;
    float next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return static_cast<float>(z >> 40) * (1.0f / 16777216.0f);
    }
};

// Random moments over the compact region (ghost planes included in jet mode). Three zones:
// x<3: rr == rb exactly (psi = 0, colour gradient 0 -> the |F| <= 1e-4 branch);
// x>=NX-4: pure red (rb = 0 -> interface indicator 0); elsewhere a smooth random interface.
void init_random(HostState &s, uint64_t seed) {
    Rng r{seed};
    for (label_t z = 1; z + 1 < NZ; ++z)
        for (label_t y = y_lo(); y <= y_hi(); ++y)
            for (label_t x = 1; x + 1 < NX; ++x) {
                const label_t id = idx(x, y, z);
                const double phi_s = std::tanh(0.9 * (std::sin(0.7 * x + 0.3 * y) +
                                                      std::cos(0.5 * z - 0.4 * y)) +
                                               0.3 * (r.next() - 0.5));
                const double rho = 1.0 + 0.1 * (r.next() - 0.5);
                double rr = 0.5 * (1.0 + phi_s) * rho, rb = 0.5 * (1.0 - phi_s) * rho;
                if (x < 3) { rr = 0.5 * rho; rb = 0.5 * rho; }
                if (x + 4 >= NX) { rr = rho; rb = 0.0; }
                s.buf[0][id] = static_cast<float>(rr);
                s.buf[1][id] = static_cast<float>(rb);
                for (int f = 2; f < 5; ++f) s.buf[f][id] = 0.1f * (r.next() - 0.5f);
                for (int f = 5; f < 11; ++f) s.buf[f][id] = 2.0e-3f * (r.next() - 0.5f);
            }
}

// Diffuse drop (blue inside): phi = tanh(2(R - r)/w), rr = (1-phi)/2, rb = (1+phi)/2, computed in
// double and rounded once; centre = bubble_x0/y0/z0 of constants.cuh.
void init_bubble_tanh(HostState &s, double width) {
    for (label_t z = 1; z + 1 < NZ; ++z)
        for (label_t y = 1; y + 1 < NY; ++y)
            for (label_t x = 1; x + 1 < NX; ++x) {
                const label_t id = idx(x, y, z);
                const double dx = x - static_cast<double>(bubble_x0);
                const double dy = y - static_cast<double>(bubble_y0);
                const double dz = z - static_cast<double>(bubble_z0);
                const double rr_ = std::sqrt(dx * dx + dy * dy + dz * dz);
                const double phi = std::tanh(2.0 * (static_cast<double>(bubble_radius) - rr_) / width);
                s.buf[0][id] = static_cast<float>(0.5 * (1.0 - phi));
                s.buf[1][id] = static_cast<float>(0.5 * (1.0 + phi));
            }
}

void jet_bc(HostState &B, const HostState &Aold) {
    // JetBoundaryConditions (kernels.cu:102-114): interior x,z only.
    for (label_t z = 0; z < NZ; ++z)
        for (label_t x = 0; x < NX; ++x) {
            if (inlet_outlet_interior(x, z)) continue;
            inlet_bc_calculation(B.view, x, z);
            outlet_neumann_bc_calculation(B.view, Aold.view, x, z);
        }
}

void rcs_all(const HostState &A, HostState &B) {
    // RCSKernel (kernels.cu:119-131): every non-ghost cell.
#pragma omp parallel for schedule(static)
    for (long zz = 0; zz < static_cast<long>(NZ); ++zz) {
        const label_t z = static_cast<label_t>(zz);
        for (label_t y = 0; y < NY; ++y)
            for (label_t x = 0; x < NX; ++x) {
                if (interior(x, y, z)) continue;
                RCS(A.view, B.view, x, y, z);
            }
    }
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s <jet|random|bubble_tanh|bubble_sharp|rti_init> <seed|width> "
                             "<steps comma list> <out prefix>\n", argv[0]);
        return 1;
    }
    const std::string mode = argv[1];
    const double arg = std::atof(argv[2]);
    std::vector<int> dumps;
    for (const char *p = argv[3]; *p;) {
        dumps.push_back(std::atoi(p));
        while (*p && *p != ',') ++p;
        if (*p == ',') ++p;
    }
    const std::string prefix = argv[4];

    if (mode.rfind("time_", 0) == 0) {
        // CPU-baseline timing: the initial state of the case (time_bubble_tanh: diffuse drop of
        // width argv[2], periodic-y build; time_jet: launch_InitJet + main.cu:26; time_rti: the
        // rtiIn.cuh profile, walls held), one untimed step, then argv[3] timed steps of the
        // reference's step (RCS over every interior cell, plus the jet boundary kernel for the
        // jet); prints seconds per step.
        const std::string what = mode.substr(5);
        const bool ok = what == "bubble_tanh" ? g_periodic_y : (what == "jet" || what == "rti") && !g_periodic_y;
        if (!ok) { std::fprintf(stderr, "%s needs the %s build\n", mode.c_str(), g_periodic_y ? "jet" : "periodic-y"); return 1; }
        HostState *A = new HostState, *B = new HostState;
        if (what == "bubble_tanh") {
            init_bubble_tanh(*A, arg);
        } else if (what == "jet") {
            for (label_t z = 0; z < NZ; ++z)
                for (label_t y = 0; y < NY; ++y)
                    for (label_t x = 0; x < NX; ++x) init_jet_bulk(A->view, x, y, z);
            for (label_t z = 0; z < NZ; ++z)
                for (label_t x = 0; x < NX; ++x)
                    if (!inlet_outlet_interior(x, z)) init_jet_inlet(A->view, x, z);
            jet_bc(*A, *A);
        } else {
            for (label_t z = 0; z < NZ; ++z)
                for (label_t y = 0; y < NY; ++y)
                    for (label_t x = 0; x < NX; ++x) init_rti_domain(A->view, x, y, z);
            // the wall planes, which the step never writes, hold in both buffers
            for (int f = 0; f < 11; ++f) std::copy(A->buf[f].begin(), A->buf[f].end(), B->buf[f].begin());
        }
        auto one = [&]() {
            rcs_all(*A, *B);
            if (what == "jet") jet_bc(*B, *A);
            std::swap(A, B);
        };
        one();
        const int steps = dumps.empty() ? 1 : dumps[0];
        const auto t0 = std::chrono::steady_clock::now();
        for (int k = 0; k < steps; ++k) one();
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("seconds_per_step %.9f cells %llu\n", dt / steps,
                    static_cast<unsigned long long>(NX - 2) * (NY - 2) * (NZ - 2));
        delete A;
        delete B;
        return 0;
    }

    HostState *A = new HostState, *B = new HostState;
    if (mode == "jet") {
        if (g_periodic_y) { std::fprintf(stderr, "jet needs the y ghost planes\n"); return 1; }
        // launch_InitJet (launch.cuh:52-72): bulk over every cell, inlet over interior x,z,
        // then main.cu:26 applies the boundary kernel with B == Aold == A.
        for (label_t z = 0; z < NZ; ++z)
            for (label_t y = 0; y < NY; ++y)
                for (label_t x = 0; x < NX; ++x) init_jet_bulk(A->view, x, y, z);
        for (label_t z = 0; z < NZ; ++z)
            for (label_t x = 0; x < NX; ++x)
                if (!inlet_outlet_interior(x, z)) init_jet_inlet(A->view, x, z);
        jet_bc(*A, *A);
    } else if (mode == "random") {
        init_random(*A, static_cast<uint64_t>(arg));
    } else if (mode == "bubble_tanh") {
        if (!g_periodic_y) { std::fprintf(stderr, "bubble needs the periodic-y build\n"); return 1; }
        init_bubble_tanh(*A, arg);
    } else if (mode == "bubble_sharp") {
        if (!g_periodic_y) { std::fprintf(stderr, "bubble needs the periodic-y build\n"); return 1; }
        for (label_t z = 1; z + 1 < NZ; ++z)
            for (label_t y = 1; y + 1 < NY; ++y)
                for (label_t x = 1; x + 1 < NX; ++x) init_static_bubble(A->view, x, y, z);
    } else if (mode == "rti_init") {
        for (label_t z = 0; z < NZ; ++z)
            for (label_t y = 0; y < NY; ++y)
                for (label_t x = 0; x < NX; ++x) init_rti_domain(A->view, x, y, z);
        dump(*A, prefix + "_step0.bin");
        return 0;
    } else {
        std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
        return 1;
    }
    dump(*A, prefix + "_step0.bin");

    int last = 0;
    for (int d : dumps) last = d > last ? d : last;
    size_t di = 0;
    for (int step = 1; step <= last; ++step) {
        rcs_all(*A, *B);
        if (!g_periodic_y) jet_bc(*B, *A);
        std::swap(A, B);
        while (di < dumps.size() && dumps[di] == step) {
            dump(*A, prefix + "_step" + std::to_string(step) + ".bin");
            ++di;
        }
    }
    std::printf("Q=%u NX=%u NY=%u NZ=%u steps=%d\n", Q, NX, NY, NZ, last);
    delete A;
    delete B;
    return 0;
}
