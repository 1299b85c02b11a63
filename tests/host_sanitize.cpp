// Host shell under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only;
// built and run by tests/test_host_sanitize.py).  Drives the C-ABI of
// include/howmhd_host.h through malformed inputs, every registered problem
// with both writers, checkpoint round trips and damaged files, and indicator
// weights files, and checks each return code.
//
//   host_sanitize <scratch dir> <weights file>
#include "howmhd_host.h"

#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

static int g_fail = 0;

static void expect(bool ok_expected, int rc, const std::string& what, const char* err) {
    if ((rc == 0) != ok_expected) {
        std::printf("FAIL %s: rc=%d (%s)\n", what.c_str(), rc, err);
        ++g_fail;
    }
}

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    const std::string dir = argv[1], weights = argv[2];
    char err[512];
    howmhd_run run;
    const std::string ok = "nx 4\nny 4\nnz 4\nnt 1\ntend 1\n";
    const std::string longtok(300, 'a');
    struct Case { std::string text; bool parses; };
    const Case cases[] = {
        {"nx 0\nny 4\nnz 4\nnt 1\ntend 1\n", false},
        {"nx 8\nny 4\nnz 4\nnt -1\ntend 1\n", false},
        {"nx 8\nny 4\nnz 4\nnt 1\ntend 0\n", false},
        {"nx abc\n", false},
        {"nx 8\nny\n", false},
        {"", false},
        {"# only a comment", false},
        {ok + "x1bc weird\n", false},
        {ok + "precision fp128\n", false},
        // over-long strings are cut to their fields, not overflowed
        {"problem " + longtok + "\n" + ok, true},
        {"x1bc " + longtok + "\n" + ok, false},
        {"indicator_weights " + longtok + "\n" + "restart " + longtok + "\n" + ok, true},
        {"unknown_key 5 6 7\n" + ok, true},
    };
    for (const Case& c : cases) {
        std::memset(&run, 0, sizeof(run));
        int rc = howmhd_host_parse(c.text.c_str(), &run, err, sizeof(err));
        howmhd_params p;
        if (rc == 0) rc = howmhd_host_gpu_params(&run, &p, err, sizeof(err));
        expect(c.parses, rc, "parse: " + c.text.substr(0, 40), err);
        char e1[1], e8[8];
        howmhd_host_parse(c.text.c_str(), &run, e1, 1);
        howmhd_host_parse(c.text.c_str(), &run, e8, 8);
        howmhd_host_parse(c.text.c_str(), &run, nullptr, 0);
    }
    // an unknown problem parses and fails at initialisation
    {
        std::memset(&run, 0, sizeof(run));
        expect(true, howmhd_host_parse(("problem nosuch\n" + ok).c_str(), &run, err, sizeof(err)), "parse nosuch", err);
        std::vector<double> q(8 * 10 * 10 * 10, 0.0);
        expect(false, howmhd_host_init_problem(&run, q.data(), err, sizeof(err)), "init nosuch", err);
    }
    // every registered problem on a ragged grid, both writers, checkpoints
    std::stringstream names(howmhd_host_problem_names());
    std::string name;
    const std::string ck = dir + "/ck.bin";
    while (std::getline(names, name)) {
        if (name.empty()) continue;
        for (const char* fmt : {"vts", "dat", "none"}) {
            const std::string in = "problem " + name + "\nnx 9\nny 7\nnz 5\nnt 1\ntend 1\nseed 3\noutput_format " + fmt + "\n";
            std::memset(&run, 0, sizeof(run));
            const int rc = howmhd_host_parse(in.c_str(), &run, err, sizeof(err));
            expect(true, rc, "parse " + name, err);
            if (rc) continue;
            howmhd_params p;
            expect(true, howmhd_host_gpu_params(&run, &p, err, sizeof(err)), "gpu_params " + name, err);
            const size_t n6 = (size_t)(run.nx + 6) * (run.ny + 6) * (run.nz + 6);
            std::vector<double> q(8 * n6, 0.0), b(3 * n6, 0.5);
            expect(true, howmhd_host_init_problem(&run, q.data(), err, sizeof(err)), "init " + name, err);
            expect(true, howmhd_host_output(&run, q.data(), 3, 0.25, dir.c_str(), err, sizeof(err)),
                   std::string("output ") + fmt, err);
            expect(true, howmhd_host_write_checkpoint(&run, ck.c_str(), q.data(), b.data(), 3, 17, 0.25, err, sizeof(err)),
                   "write_checkpoint", err);
            std::vector<double> q2(8 * n6), b2(3 * n6);
            int32_t os = 0;
            int64_t ts = 0;
            double tm = 0.0;
            expect(true, howmhd_host_read_checkpoint(&run, ck.c_str(), q2.data(), b2.data(), &os, &ts, &tm, err, sizeof(err)),
                   "read_checkpoint", err);
            if (q2 != q || b2 != b || os != 3 || ts != 17 || tm != 0.25) {
                std::printf("FAIL checkpoint round trip %s\n", name.c_str());
                ++g_fail;
            }
        }
    }
    // damaged checkpoints and weights files
    {
        std::memset(&run, 0, sizeof(run));
        howmhd_host_parse(ok.c_str(), &run, err, sizeof(err));
        const size_t n6 = (size_t)(run.nx + 6) * (run.ny + 6) * (run.nz + 6);
        std::vector<double> q(8 * n6), b(3 * n6);
        int32_t os;
        int64_t ts;
        double tm;
        const std::string trunc = dir + "/trunc.bin", text = dir + "/text.bin";
        FILE* f = std::fopen(trunc.c_str(), "wb");
        std::fwrite("HOWMHDCK\x01\0\0\0", 1, 12, f);
        std::fclose(f);
        f = std::fopen(text.c_str(), "w");
        std::fputs("not a checkpoint at all, but long enough to fill a header of 48 bytes\n", f);
        std::fclose(f);
        for (const std::string& path : {trunc, text, dir + "/none.bin"})
            expect(false, howmhd_host_read_checkpoint(&run, path.c_str(), q.data(), b.data(), &os, &ts, &tm, err, sizeof(err)),
                   "damaged checkpoint " + path, err);
        std::vector<float> P(1697);
        float thr;
        const std::string w1 = dir + "/w1.txt", w2 = dir + "/w2.txt";
        f = std::fopen(w1.c_str(), "w");
        std::fputs("# c\nhowmhd-mlp 1697 0.5\n1 2 3\n", f);
        std::fclose(f);
        f = std::fopen(w2.c_str(), "w");
        std::fputs("garbage\n", f);
        std::fclose(f);
        for (const std::string& path : {w1, w2, dir + "/nope.txt"})
            expect(false, howmhd_host_load_indicator(path.c_str(), P.data(), 1697, &thr, err, sizeof(err)),
                   "weights " + path, err);
        expect(true, howmhd_host_load_indicator(weights.c_str(), P.data(), 1697, &thr, err, sizeof(err)),
               "committed weights", err);
    }
    std::printf("host_sanitize: %d failures\n", g_fail);
    return g_fail ? 1 : 0;
}
