// how_mhd — drop-in driver for `./bin/how_mhd_mpi input.in` at one rank.
//
// Restates src/main.cpp:25-248: read the input (argv[1], or stdin), finalise
// the parameters, initialise the problem on the host, hand q(0) to the GPU
// library, build the face field, write output 0, then for each output
// interval run CFL-limited hydro steps (dt clipped to the next output time,
// abort on dt <= 0 or a non-finite time) and write the output.  Added:
// `max_steps` > 0 stops after that many hydro steps (`nstep` is parsed and
// not used, as in the reference), a closing [timing] line reports
// zone-cycles/s of the step loop, `checkpoint_every N` writes
// checkpoint/ckptNNNNNN.bin (q(0), faces, output step, steps, time) after
// every N-th output, and `restart <file>` resumes from such a file with the
// next output interval.  Without the learned indicator a resumed run
// continues the uninterrupted one bit for bit; with it, the plain-WENO
// fallback state is not in the checkpoint and starts afresh, so the resumed
// run may route steps the uninterrupted run sent to the plain sweep.
#include "howmhd_gpu.h"
#include "howmhd_host.h"
#include "howmhd_mlp.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace {

void print_banner() {
    std::cout << "\n"
              << "  HOW-MHD  |  WENO5  |  SSPRK(5,4)  |  4th-order CT  |  MI355X (gfx950)\n"
              << "  High-order WENO magnetohydrodynamics, single-GPU HIP time step\n"
              << "  Method: Seo & Ryu (2023), ApJ 953, 39\n\n";
}

// print_log, src/main.cpp:275-341 (one rank)
void print_log(const howmhd_run& par) {
    std::cout << "\n";
    std::cout << "===== LOG rank 0 =====\n";
    std::cout << "MPI ranks          = 1\n";
    std::cout << "MPI layout         = 1 x 1 x 1\n";
    std::cout << "rank coords        = (0, 0, 0)\n";
    std::cout << "global offset      = 0 0 0\n";
    std::cout << "global nx ny nz    = " << par.gnx << " " << par.gny << " " << par.gnz << "\n";
    std::cout << "local  nx ny nz    = " << par.nx << " " << par.ny << " " << par.nz << "\n";
    std::cout << "nt                 = " << par.nt << "\n";
    std::cout << "xsize ysize zsize  = " << par.xsize << " " << par.ysize << " " << par.zsize << "\n";
    std::cout << "tend               = " << par.tend << "\n";
    std::cout << "dx dy dz           = " << par.dx << " " << par.dy << " " << par.dz << "\n";
    std::cout << "boundary           = " << par.x1bc << " " << par.x2bc << " " << par.x3bc << "\n";
    std::cout << "output_format      = " << par.output_format << "\n";
    std::cout << "precision          = " << par.precision << "\n";
    std::cout << "indicator          = " << par.indicator << "\n";
    std::cout << "============================\n";
}

[[noreturn]] void die(const std::string& msg) {
    std::cerr << msg << "\n";
    std::abort();
}

void check(int rc, howmhd_gpu_ctx* ctx, const char* what) {
    if (rc != HOWMHD_OK) die(std::string("Error: ") + what + ": " + howmhd_gpu_last_error(ctx));
}

}  // namespace

int main(int argc, char* argv[]) {
    std::cout.setf(std::ios::unitbuf);
    std::cerr.setf(std::ios::unitbuf);

    std::string text;
    if (argc > 1) {
        std::ifstream f(argv[1]);
        if (!f) die(std::string("Error: cannot open input file: ") + argv[1]);
        std::stringstream ss;
        ss << f.rdbuf();
        text = ss.str();
    } else {
        std::stringstream ss;
        ss << std::cin.rdbuf();
        text = ss.str();
    }

    char err[4096] = {0};
    howmhd_run par;
    if (howmhd_host_parse(text.c_str(), &par, err, sizeof(err)) != 0) die(err);
    if (par.px > 1 || par.py > 1 || par.pz > 1)
        std::cerr << "Warning: px/py/pz were provided, but this program runs one rank on one GPU. "
                  << "Ignoring MPI layout and using one rank.\n";

    print_banner();
    print_log(par);

    howmhd_params gp;
    if (howmhd_host_gpu_params(&par, &gp, err, sizeof(err)) != 0) die(err);

    const size_t ncell = (size_t)(par.nx + 6) * (par.ny + 6) * (par.nz + 6);
    std::vector<double> q0(8 * ncell, 0.0), faces;
    const bool resume = par.restart[0] != '\0';
    int32_t start_out = 0;
    int64_t steps_done = 0;
    double time = 0.0;
    if (resume) {
        faces.assign(3 * ncell, 0.0);
        if (howmhd_host_read_checkpoint(&par, par.restart, q0.data(), faces.data(), &start_out, &steps_done, &time, err,
                                        sizeof(err)) != 0)
            die(err);
        std::cout << "restart            = " << par.restart << " (output step " << start_out << ", time " << time
                  << ", " << steps_done << " hydro steps)\n";
    } else if (howmhd_host_init_problem(&par, q0.data(), err, sizeof(err)) != 0) {
        die(err);
    }
    // reject an unknown output format before any step runs (src/output.cpp:565-569)
    {
        const std::string f = par.output_format;
        if (f != "dat" && f != "vtk" && f != "vts" && f != "none")
            die("Error: unknown output_format = " + f + "\nAvailable options: dat, vtk, none");
    }

    howmhd_gpu_ctx* ctx = nullptr;
    if (howmhd_gpu_create(&gp, &ctx) != 0) die(std::string("Error: ") + howmhd_gpu_last_error(nullptr));
    if (gp.indicator == HOWMHD_INDICATOR_MLP) {
        if (par.indicator_weights[0] == '\0') die("Error: indicator mlp needs indicator_weights <file>");
        std::vector<float> w(HOWMHD_MLP_NPARAM);
        float thr = 0.5f;
        if (howmhd_host_load_indicator(par.indicator_weights, w.data(), HOWMHD_MLP_NPARAM, &thr, err, sizeof(err)) != 0)
            die(err);
        check(howmhd_gpu_set_indicator(ctx, w.data(), thr, 1), ctx, "set_indicator");
        std::cout << "indicator weights  = " << par.indicator_weights << " (threshold " << thr << ")\n";
    }
    check(howmhd_gpu_upload_q0(ctx, q0.data()), ctx, "upload_q0");
    if (resume) check(howmhd_gpu_upload_faces(ctx, faces.data()), ctx, "upload_faces");
    else check(howmhd_gpu_init_faces(ctx), ctx, "init_faces");

    const double dtout = par.tend / static_cast<double>(par.nt);
    std::cout << "dtout = " << dtout << std::endl;

    auto write_output = [&](int step, double t) {
        if (std::string(par.output_format) == "none") return;
        check(howmhd_gpu_download_q0(ctx, q0.data()), ctx, "download_q0");
        if (howmhd_host_output(&par, q0.data(), step, t, "", err, sizeof(err)) != 0) die(err);
    };

    auto write_checkpoint = [&](int out_step, long steps, double t) {
        char name[64];
        std::snprintf(name, sizeof(name), "checkpoint/ckpt%06d.bin", out_step);
        std::filesystem::create_directories("checkpoint");
        faces.resize(3 * ncell);
        check(howmhd_gpu_download_q0(ctx, q0.data()), ctx, "download_q0");
        check(howmhd_gpu_download_faces(ctx, faces.data()), ctx, "download_faces");
        if (howmhd_host_write_checkpoint(&par, name, q0.data(), faces.data(), out_step, steps, t, err, sizeof(err)) != 0)
            die(err);
        std::cout << "checkpoint = " << name << std::endl;
    };

    if (!resume) {
        write_output(0, time);
        std::cout << "output step = 0 time = " << time << std::endl;
    }

    long total_steps = (long)steps_done;
    double loop_seconds = 0.0;
    bool stop = false;
    for (int out_step = start_out + 1; out_step <= par.nt && !stop; ++out_step) {
        const double target_time = dtout * static_cast<double>(out_step);
        int hydro_step = 0;
        double dt = 0.0;
        const auto t0 = std::chrono::steady_clock::now();
        while (time < target_time) {
            if (par.max_steps > 0 && total_steps >= par.max_steps) { stop = true; break; }
            check(howmhd_gpu_tstep(ctx, &dt), ctx, "tstep");
            if (time + dt > target_time) dt = target_time - time;
            if (dt <= 0.0) {
                std::cerr << "Error: non-positive dt detected.\n";
                std::cerr << "time        = " << time << "\n";
                std::cerr << "target_time = " << target_time << "\n";
                std::cerr << "dt          = " << dt << "\n";
                break;
            }
            if (!par.quiet)
                std::cout << "[advance begin] out_step = " << out_step << " hydro_step = " << hydro_step + 1
                          << " time = " << time << " dt = " << dt << " target_time = " << target_time << std::endl;
            check(howmhd_gpu_ssprk(ctx, dt), ctx, "ssprk");
            time += dt;
            hydro_step++;
            total_steps++;
            if (!par.quiet)
                std::cout << "[advance done ] out_step = " << out_step << " hydro_step = " << hydro_step
                          << " time = " << time << " dt = " << dt << std::endl;
            if (!std::isfinite(time)) {
                std::cerr << "Error: time became non-finite.\n";
                break;
            }
            if (hydro_step % 50 == 0)
                std::cout << "hydro_step = " << hydro_step << " time = " << time << " dt = " << dt
                          << " target_time = " << target_time << std::endl;
        }
        check(howmhd_gpu_synchronize(ctx), ctx, "synchronize");
        loop_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        write_output(out_step, time);
        std::cout << "output step = " << out_step << " time = " << time << " last_dt = " << dt
                  << " hydro_steps = " << hydro_step << std::endl;
        if (!stop && par.checkpoint_every > 0 && out_step % par.checkpoint_every == 0)
            write_checkpoint(out_step, total_steps, time);
    }
    const double zones = (double)par.gnx * par.gny * par.gnz;
    const long run_steps = total_steps - (long)steps_done;  // steps of this invocation
    std::cout << "[timing] hydro_steps = " << run_steps << " loop_seconds = " << loop_seconds
              << " zone_cycles_per_s = " << (loop_seconds > 0 ? zones * run_steps / loop_seconds : 0.0)
              << std::endl;
    howmhd_gpu_destroy(ctx);
    return 0;
}
