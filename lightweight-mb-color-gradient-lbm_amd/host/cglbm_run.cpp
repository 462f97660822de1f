// Host driver over the C ABI in the shape of the reference's main program (main.cu:13-69):
// allocate the state, initialise the case (plus the boundary planes), then steps 1..NSTEP-1 with
// an MLUPS clock that pauses for output, a VTI file every NOUTPUT steps (and at step 0), and the
// MLUPS summary at the end. Configuration is runtime (command line) instead of constants.cuh.
//
// Output files are named and laid out as io/vtiWriter.cuh:43-54 and 92-145 write them
// (<out>/jet_%08d.vti). For the jet case the default extent is the reference's full grid, ghost
// layers included: the x/z ghost shell holds the jet's bulk initial state (rhor = 1, the rest 0),
// which the reference never updates (kernels.cu:125-128), so the files are byte-identical to the
// reference program's for the same configuration.
//
//   cglbm_run [--case jet|bubble|rti] [--q 19|27] [--nx N] [--ny N] [--nz N] [--nstep N]
//             [--noutput N] [--tau T] [--sigma S] [--beta B] [--jet-radius R] [--jet-velocity U]
//             [--jet-noise E] [--seed S] [--zero-gradient] [--radius R] [--width W]
//             [--gravity G] [--atwood A] [--storage f32|f16] [--extent reference|stored]
//             [--out DIR]
// Defaults: constants.cuh as shipped (128^3 grid with ghosts = 126^3 interior, D3Q27, jet R 12,
// U 0.05, tau 1, sigma 0.1, beta 0.9), NSTEP 1001, NOUTPUT 1000, out JET_VTK/cglbm.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cglbm.h"

namespace {

[[noreturn]] void die(const char *what, int rc) {
    std::fprintf(stderr, "cglbm_run: %s failed (%d): %s\n", what, rc, cglbm_last_error());
    std::exit(1);
}

void check(int rc, const char *what) {
    if (rc != CGLBM_OK) die(what, rc);
}

struct Options {
    std::string kind = "jet";
    int q = 27, nx = 126, ny = 126, nz = 126;
    int nstep = 1001, noutput = 1000;
    float tau = 1.0f, sigma = 0.1f, beta = 0.9f;
    float jet_radius = 12.0f, jet_velocity = 0.05f, jet_noise = 0.0f;
    uint32_t seed = 0;
    bool zero_gradient = false;
    double radius = 30.0, width = 4.0;
    float gravity = 0.0f, atwood = 0.0f;
    int storage = CGLBM_STORE_F32;
    bool reference_extent = true;
    std::string out = "JET_VTK/cglbm";
};

Options parse(int argc, char **argv) {
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto val = [&]() -> const char * {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "cglbm_run: %s needs a value\n", a.c_str());
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--case") o.kind = val();
        else if (a == "--q") o.q = std::atoi(val());
        else if (a == "--nx") o.nx = std::atoi(val());
        else if (a == "--ny") o.ny = std::atoi(val());
        else if (a == "--nz") o.nz = std::atoi(val());
        else if (a == "--nstep") o.nstep = std::atoi(val());
        else if (a == "--noutput") o.noutput = std::atoi(val());
        else if (a == "--tau") o.tau = std::strtof(val(), nullptr);
        else if (a == "--sigma") o.sigma = std::strtof(val(), nullptr);
        else if (a == "--beta") o.beta = std::strtof(val(), nullptr);
        else if (a == "--jet-radius") o.jet_radius = std::strtof(val(), nullptr);
        else if (a == "--jet-velocity") o.jet_velocity = std::strtof(val(), nullptr);
        else if (a == "--jet-noise") o.jet_noise = std::strtof(val(), nullptr);
        else if (a == "--seed") o.seed = static_cast<uint32_t>(std::strtoul(val(), nullptr, 10));
        else if (a == "--zero-gradient") o.zero_gradient = true;
        else if (a == "--radius") o.radius = std::strtod(val(), nullptr);
        else if (a == "--width") o.width = std::strtod(val(), nullptr);
        else if (a == "--gravity") o.gravity = std::strtof(val(), nullptr);
        else if (a == "--atwood") o.atwood = std::strtof(val(), nullptr);
        else if (a == "--storage") o.storage = std::string(val()) == "f16" ? CGLBM_STORE_F16 : CGLBM_STORE_F32;
        else if (a == "--extent") o.reference_extent = std::string(val()) == "reference";
        else if (a == "--out") o.out = val();
        else {
            std::fprintf(stderr, "cglbm_run: unknown option %s\n", a.c_str());
            std::exit(2);
        }
    }
    if (o.kind != "jet" && o.kind != "bubble" && o.kind != "rti") {
        std::fprintf(stderr, "cglbm_run: --case must be jet, bubble or rti\n");
        std::exit(2);
    }
    return o;
}

// write_vti_step_device (io/vtiWriter.cuh:151-164): device state -> host -> <out>/jet_%08d.vti.
void write_step(cglbm_domain *dom, const Options &o, const cglbm_config &c, int step) {
    char name[64];
    std::snprintf(name, sizeof(name), "/jet_%08d.vti", step);
    const std::string path = o.out + name;
    if (!(o.reference_extent && o.kind == "jet")) {
        check(cglbm_write_vti(dom, path.c_str()), "cglbm_write_vti");
    } else {
        // stored extent: nx x (ny + 2) x nz; reference extent: (nx + 2) x (ny + 2) x (nz + 2)
        const int64_t sx = c.nx, sy = c.ny + 2, sz = c.nz, n = sx * sy * sz;
        const int64_t fx = c.nx + 2, fy = c.ny + 2, fz = c.nz + 2, nf = fx * fy * fz;
        std::vector<float> dev(5 * n), full(5 * nf, 0.0f);
        check(cglbm_copy_out(dom, 0x1Fu, dev.data()), "cglbm_copy_out");
        for (int64_t i = 0; i < nf; ++i) full[i] = 1.0f;  // jet bulk state in the ghost shell: rhor = 1
        for (int f = 0; f < 5; ++f)
            for (int64_t z = 0; z < sz; ++z)
                for (int64_t y = 0; y < sy; ++y)
                    std::memcpy(&full[f * nf + (z + 1) * fx * fy + y * fx + 1], &dev[f * n + z * sx * sy + y * sx],
                                sizeof(float) * sx);
        check(cglbm_write_vti_host(path.c_str(), static_cast<int>(fx), static_cast<int>(fy), static_cast<int>(fz),
                                   &full[0], &full[nf], &full[2 * nf], &full[3 * nf], &full[4 * nf]),
              "cglbm_write_vti_host");
    }
    std::printf("Wrote VTI file: \"%s\"\n", path.c_str());
}

}  // namespace

int main(int argc, char **argv) {
    const Options o = parse(argc, argv);
    cglbm_config c{};
    c.nx = o.nx;
    c.ny = o.ny;
    c.nz = o.nz;
    c.q = o.q;
    c.ybc = o.kind == "jet" ? CGLBM_Y_JET : (o.kind == "rti" ? CGLBM_Y_WALLS : CGLBM_Y_PERIODIC);
    c.outlet = o.zero_gradient ? CGLBM_OUTLET_ZERO_GRADIENT : CGLBM_OUTLET_CONVECTIVE;
    c.tau = o.tau;
    c.sigma = o.sigma;
    c.beta = o.beta;
    c.jet_x0 = static_cast<float>(o.nx - 1) / 2.0f;
    c.jet_z0 = static_cast<float>(o.nz - 1) / 2.0f;
    c.jet_radius = o.jet_radius;
    c.jet_velocity = o.jet_velocity;
    c.jet_noise = o.jet_noise;
    c.seed = o.seed;
    c.gravity = o.gravity;
    c.atwood = o.atwood;

    cglbm_domain *dom = nullptr;
    check(cglbm_domain_create(&c, o.storage, &dom), "cglbm_domain_create");  // allocate_device_memory
    if (o.kind == "jet") check(cglbm_init_jet(dom), "cglbm_init_jet");   // launch_InitJet + boundary planes
    else if (o.kind == "bubble")
        check(cglbm_init_bubble(dom, (o.nx - 1) / 2.0, (o.ny - 1) / 2.0, (o.nz - 1) / 2.0, o.radius, o.width, 0),
              "cglbm_init_bubble");
    else check(cglbm_init_rti(dom, 2.0, o.width, 8, 0.2, o.seed), "cglbm_init_rti");

    const int final_step = o.nstep - 1;
    const double cells = static_cast<double>(o.nx) * o.ny * o.nz;
    double ms_total = 0.0;
    int64_t steps_timed = 0;
    if (o.noutput > 0) write_step(dom, o, c, 0);
    int step = 1;
    while (step <= final_step) {
        // advance to the next output step (or the end) in one timed run; output is off the clock
        int next = final_step + 1;
        if (o.noutput > 0) next = std::min(final_step + 1, (step + o.noutput - 1) / o.noutput * o.noutput + 1);
        const int n = next - step;
        double ms = 0.0;
        check(cglbm_time_steps(dom, n, &ms), "cglbm_time_steps");
        ms_total += ms;
        steps_timed += n;
        step = next;
        const int done = step - 1;
        std::printf("step %d / %d  %.1f MLUPS\n", done, final_step, cells * steps_timed / (ms_total * 1e3));
        if (o.noutput > 0 && done % o.noutput == 0) write_step(dom, o, c, done);
    }
    double diag[6];
    check(cglbm_reduce(dom, diag), "cglbm_reduce");
    std::printf("MLUPS %.1f over %lld steps of %.0f cells (%.3f ms)\n", cells * steps_timed / (ms_total * 1e3),
                static_cast<long long>(steps_timed), cells, ms_total);
    std::printf("mass red %.9e blue %.9e  non-finite cells %.0f\n", diag[0], diag[1], diag[5]);
    check(cglbm_domain_destroy(dom), "cglbm_domain_destroy");
    return 0;
}
