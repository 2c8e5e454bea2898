// cuvegas_main: the standalone VEGAS+ program of the reference (src/main.cu:277-798)
// on the MI355X library, driven by the same configuration header.
//
// The reference compiles one executable per configuration: src/config.cuh
// fixes the settings (tot_eval, max_it, skip, n_intervals, alpha, beta, n_dim)
// and the integrand (__host__ __device__ double integrand(double* x), usually a
// call into its library, src/commons/functions.cu), and main() runs the loop on
// the GPUs and prints the result.  Here one program reads such a header at run
// time: the settings from its initialisers, the integrand compiled by hipRTC
// through the C ABI (include/cuvegas.h), with the library's functions
// (vg_functions.h) standing in for "commons/functions.cuh".  It prints the
// reference's lines (main.cu:292,428-430,791-797) followed by this build's
// extra figures.  The reference's sizing is kept: n_eval_it = tot_eval /
// max_it, n_strat = (int)pow(n_eval_it / 4, 1 / n_dim) (main.cu:305,368), and
// its fixed seed 123 (main.cu:16, here the Philox key).
//
// Usage: cuvegas_main CONFIG [--precision fp64|fp32] [--seed N] [--eager]
//                            [--compile-only]
//   --precision fp32 needs an integrand written in vg_real (the header's
//   double form is the FP64 form); --eager launches each iteration instead of
//   replaying a hipGraph; --compile-only parses and compiles the integrand
//   (no GPU needed) and prints the settings.
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "../../include/cuvegas.h"
#include "vg_main_embed.inc"  // kVgFunctionsSrc: vg_functions.h as a string (written by build.py)

namespace {

double now_ms() {
    using clk = std::chrono::steady_clock;
    return std::chrono::duration<double, std::milli>(clk::now().time_since_epoch()).count();
}

// The text without C and C++ comments (string literals are not expected in a
// configuration header).
std::string strip_comments(const std::string& s) {
    std::string out;
    out.reserve(s.size());
    for (size_t i = 0; i < s.size();) {
        if (s.compare(i, 2, "//") == 0) {
            while (i < s.size() && s[i] != '\n') ++i;
        } else if (s.compare(i, 2, "/*") == 0) {
            const size_t e = s.find("*/", i + 2);
            i = e == std::string::npos ? s.size() : e + 2;
            out += ' ';
        } else {
            out += s[i++];
        }
    }
    return out;
}

bool ident_char(char c) { return std::isalnum((unsigned char)c) || c == '_'; }

// Value of the initialiser "key = <number>;" (the first occurrence of key as a
// whole word followed by '='): 1 with *out set, 0 if there is none, -1 if the
// initialiser is not a plain number (an expression, which this program does
// not evaluate).
int find_setting(const std::string& text, const char* key, double* out) {
    const size_t n = std::strlen(key);
    for (size_t p = text.find(key); p != std::string::npos; p = text.find(key, p + 1)) {
        if (p > 0 && ident_char(text[p - 1])) continue;
        if (p + n < text.size() && ident_char(text[p + n])) continue;
        size_t q = p + n;
        while (q < text.size() && std::isspace((unsigned char)text[q])) ++q;
        if (q >= text.size() || text[q] != '=' || (q + 1 < text.size() && text[q + 1] == '=')) continue;
        const char* start = text.c_str() + q + 1;
        char* end = nullptr;
        const double v = std::strtod(start, &end);
        if (end == start) return -1;
        while (*end && std::isspace((unsigned char)*end)) ++end;
        while (*end == 'l' || *end == 'L' || *end == 'u' || *end == 'U' || *end == 'f' || *end == 'F') ++end;
        while (*end && std::isspace((unsigned char)*end)) ++end;
        if (*end != ';') return -1;
        *out = v;
        return 1;
    }
    return 0;
}

// The configuration as hipRTC source: an include of the integrand library is
// replaced by the library, includes of the C maths and I/O headers are dropped
// (hipRTC provides the maths functions), any other include is refused.
bool integrand_source(const std::string& cfg, std::string* src, std::string* err) {
    std::istringstream in(cfg);
    std::string line, out;
    bool have_lib = false;
    while (std::getline(in, line)) {
        size_t a = 0;
        while (a < line.size() && std::isspace((unsigned char)line[a])) ++a;
        if (line.compare(a, 8, "#include") != 0) {
            out += line + "\n";
            continue;
        }
        if (line.find("functions.cuh") != std::string::npos) {
            if (!have_lib) out += kVgFunctionsSrc;
            have_lib = true;
        } else if (line.find("math.h") != std::string::npos || line.find("<cmath>") != std::string::npos ||
                   line.find("stdio.h") != std::string::npos || line.find("<cstdio>") != std::string::npos) {
            // nothing: provided by hipRTC, or not used by device code
        } else {
            *err = "unsupported include in the configuration: " + line.substr(a);
            return false;
        }
        out += "\n";  // keep line numbers of hipRTC messages close to the file's
    }
    *src = out;
    return true;
}

int usage() {
    std::fprintf(stderr,
                 "usage: cuvegas_main CONFIG [--precision fp64|fp32] [--seed N] [--eager] [--compile-only]\n");
    return 2;
}

}  // namespace

int main(int argc, char** argv) {
    const double t_start = now_ms();
    const char* path = nullptr;
    int precision = CVG_FP64, use_graph = 1, compile_only = 0;
    unsigned seed = 123u;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--precision") && i + 1 < argc) {
            const char* v = argv[++i];
            if (!std::strcmp(v, "fp64")) precision = CVG_FP64;
            else if (!std::strcmp(v, "fp32")) precision = CVG_FP32C;
            else return usage();
        } else if (!std::strcmp(argv[i], "--seed") && i + 1 < argc) {
            seed = (unsigned)std::strtoul(argv[++i], nullptr, 10);
        } else if (!std::strcmp(argv[i], "--eager")) {
            use_graph = 0;
        } else if (!std::strcmp(argv[i], "--compile-only")) {
            compile_only = 1;
        } else if (argv[i][0] == '-' || path) {
            return usage();
        } else {
            path = argv[i];
        }
    }
    if (!path) return usage();
    std::ifstream f(path);
    if (!f) {
        std::fprintf(stderr, "cannot read %s\n", path);
        return 2;
    }
    std::stringstream buf;
    buf << f.rdbuf();
    const std::string cfg = buf.str();
    const std::string text = strip_comments(cfg);

    // settings, with the values of the reference's own header where one is missing
    double tot_eval = 1e7, max_it = 20, skip = 5, n_intervals = 1024, alpha = 0.5, beta = 0.75, n_dim = 0,
           max_batch = 0;
    struct {
        const char* key;
        double* val;
    } settings[] = {{"tot_eval", &tot_eval}, {"max_it", &max_it}, {"skip", &skip},
                    {"n_intervals", &n_intervals}, {"alpha", &alpha}, {"beta", &beta},
                    {"max_batch_size", &max_batch},  // reported only: an iteration is one launch here
                    {"n_dim", &n_dim}};
    for (const auto& st : settings) {
        if (find_setting(text, st.key, st.val) < 0) {
            std::fprintf(stderr, "%s: the initialiser of %s is not a plain number\n", path, st.key);
            return 2;
        }
    }
    if (n_dim < 1 || n_dim > 64 || n_dim != std::floor(n_dim)) {
        std::fprintf(stderr, "%s: no usable 'const int n_dim = <1..64>;'\n", path);
        return 2;
    }
    if (max_it < 1 || skip < 0 || skip >= max_it || tot_eval < max_it || n_intervals < 2) {
        std::fprintf(stderr, "%s: unusable settings (tot_eval %.0f, max_it %.0f, skip %.0f, n_intervals %.0f)\n",
                     path, tot_eval, max_it, skip, n_intervals);
        return 2;
    }
    std::string src, err;
    if (!integrand_source(cfg, &src, &err)) {
        std::fprintf(stderr, "%s: %s\n", path, err.c_str());
        return 2;
    }

    cvg_params p;
    std::memset(&p, 0, sizeof p);
    p.n_dim = (int)n_dim;
    p.n_intervals = (int)n_intervals;
    p.max_it = (int)max_it;
    p.skip = (int)skip;
    p.precision = precision;
    p.n_eval_it = (long long)tot_eval / (long long)max_it;
    p.n_strat = (int)std::pow((double)p.n_eval_it / 4.0, 1.0 / n_dim);
    if (p.n_strat < 1) p.n_strat = 1;
    p.alpha = alpha;
    p.beta = beta;
    p.seed = seed;

    if (!compile_only) {
        int n_cu = 0, clock_khz = 0;
        char arch[64] = {0};
        if (cvg_device_info(&n_cu, &clock_khz, arch, sizeof arch) != CVG_OK) {
            std::fprintf(stderr, "%s\n", cvg_last_error());
            return 1;
        }
        std::printf("GPUs: 1\n");
    }
    const double t_context = now_ms();

    cvg_program* prog = nullptr;
    if (cvg_program_create(src.c_str(), &p, nullptr, &prog) != CVG_OK) {
        std::fprintf(stderr, "%s\n%s\n", cvg_last_error(), prog ? cvg_program_log(prog) : "");
        if (prog) cvg_program_destroy(prog);
        return 1;
    }
    long long cubes = 1;
    for (int i = 0; i < p.n_dim; ++i) cubes *= p.n_strat;
    std::printf("Strat: %i\n", p.n_strat);
    std::printf("Cubes: %lli\n", cubes);
    std::printf("Intervals: %i\n", p.n_intervals);
    if (compile_only) {
        std::printf("Dimensions: %i\nEvaluations per iteration: %lli\nIterations: %i (skip %i)\n"
                    "Alpha: %g\nBeta: %g\nSeed: %u\nPrecision: %s\nCompiled: %lli bytes of gfx950 code\n",
                    p.n_dim, p.n_eval_it, p.max_it, p.skip, p.alpha, p.beta, p.seed,
                    precision == CVG_FP64 ? "fp64" : "fp32", cvg_program_code_size(prog));
        cvg_program_destroy(prog);
        return 0;
    }
    cvg_state* st = nullptr;
    cvg_result r;
    int rc = cvg_state_create(prog, &p, &st);
    if (rc == CVG_OK) rc = cvg_state_run(st, p.max_it, use_graph);
    if (rc == CVG_OK) rc = cvg_state_result(st, &r);
    if (rc != CVG_OK) {
        std::fprintf(stderr, "%s\n", cvg_last_error());
        if (st) cvg_state_destroy(st);
        cvg_program_destroy(prog);
        return 1;
    }
    cvg_state_destroy(st);
    cvg_program_destroy(prog);
    const double t_end = now_ms();

    std::printf("Result: %.12f\n", r.result);
    std::printf("Error: %.12f\n", r.error);
    std::printf("Total evals: %lli\n", r.tot_nevals);
    std::printf("Time elapsed %f ms\n", t_end - t_start);
    std::printf("Iteration avg time %f ms\n", r.loop_ms / p.max_it);
    std::printf("Nocontext %f ms\n", t_end - t_context);
    std::printf("Context creation %f ms\n", t_context - t_start);
    // this build's figures, after the reference's lines
    std::printf("Chi2/dof: %.6f\n", r.chi2_dof);
    std::printf("Error (unbiased per-cube variance): %.12f\n", r.error_unbiased);
    std::printf("Chi2/dof (unbiased): %.6f\n", r.chi2_dof_unbiased);
    std::printf("Compile %f ms, setup %f ms, loop %f ms\n", r.compile_ms, r.setup_ms, r.loop_ms);
    if (max_batch > 0) std::printf("max_batch_size %.0f: not needed (one launch per iteration)\n", max_batch);
    return 0;
}
