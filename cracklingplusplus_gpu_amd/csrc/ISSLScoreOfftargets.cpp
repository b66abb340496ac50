// ISSLScoreOfftargets.cpp -- standalone scorer CLI with the reference's
// argument list and output format (ISSLScoreOfftargets/ISSLScoreOfftargets.cpp:36-647):
//
//   ISSLScoreOfftargets [issltable] [query file] [max distance]
//                       [score-threshold] [score-method] [output-file-optional]
//
// stdout: "SEQ\tMIT\tCFD" per query (%f, or -1 when not computed);
// output file: "# Query\tMIT_Score\tCFD_Score\tOfftarget_Count" then one
// line per query, Offtarget_Count being the occurrence-weighted count of
// sites within 4 mismatches; stderr: timings and the execution summary.
//
// Differences: the output file is written at the given path (the reference
// prefixes a fixed Windows directory, :562-586), or <query stem>_results.txt
// in the working directory; five arguments are required (the reference
// checks argc < 4 but reads argv[4] and argv[5]).  The distance limit is 4
// whatever the third argument says, as in the reference (:53,369).
// Scores in the ORDERED form (the reference's FP64 sums bit for bit);
// ISSL_FORM=batched selects the BATCHED form, ISSL_FORM=plain the PLAIN kernel.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <string>
#include <vector>

#include "ISSLScoringModule.hpp"
#include "cli_common.hpp"
#include "issl_gpu.h"

namespace {

double secs(std::chrono::high_resolution_clock::time_point a,
            std::chrono::high_resolution_clock::time_point b) {
    return std::chrono::duration<double>(b - a).count();
}

}  // namespace

using issl_cli::die;

int main(int argc, char **argv) {
    using clk = std::chrono::high_resolution_clock;
    const auto t_total = clk::now();
    if (argc < 6) {
        fprintf(stderr,
                "Usage: %s [issltable] [query file] [max distance] "
                "[score-threshold] [score-method] [output-file-optional]\n",
                argv[0]);
        exit(1);
    }
    const int maxDist = atoi(argv[3]);
    if (maxDist != 4) fprintf(stderr, "Note: sites within 4 mismatches are scored (max distance %d ignored)\n", maxDist);
    const double threshold = atof(argv[4]);
    int method;
    bool calcMit, calcCfd;
    issl_cli::parse_method(argv[5], &method, &calcMit, &calcCfd);

    const auto t_load = clk::now();
    issl_index *ix = nullptr;
    const char *dev = std::getenv("ISSL_DEVICE");
    int rc = issl_index_open(argv[1], dev ? atoi(dev) : 0, &ix);
    if (rc != ISSL_OK) die(rc, "Error reading index");
    issl_index_info info{};
    issl_index_get_info(ix, &info);
    const auto t_loaded = clk::now();
    printf("Index loading time: %g seconds\n", secs(t_load, t_loaded));

    const auto t_query = clk::now();
    const std::vector<uint64_t> sigs = issl_cli::read_queries(argv[2], info.seq_len + 1);
    const size_t q = sigs.size();
    printf("Query loading time: %g seconds\n", secs(t_query, clk::now()));

    const auto t_score = clk::now();
    std::vector<double> mitv(q), cfdv(q);
    std::vector<uint64_t> cnt(q);
    issl_params p{};
    p.max_dist = 4;
    p.method = method;
    p.threshold = threshold;
    p.form = issl_form_from_env();
    p.precision = ISSL_PREC_FP64;
    p.flags = ISSL_FLAG_DISTINCT;
    issl_results r{};
    r.mit = mitv.data();
    r.cfd = cfdv.data();
    r.occ_count = cnt.data();
    rc = issl_score(ix, sigs.data(), q, &p, &r);
    if (rc != ISSL_OK) die(rc, "ISSL scoring failed");
    const auto t_scored = clk::now();
    printf("ISSL scoring time: %g seconds\n", secs(t_score, t_scored));

    const auto t_out = clk::now();
    uint64_t total = 0;
    for (uint64_t c : cnt) total += c;
    uint64_t distinct = 0;
    issl_distinct_count(ix, &distinct);
    const double procMs = secs(t_query, t_scored) * 1e3;
    const double loadMs = secs(t_load, t_loaded) * 1e3;
    fprintf(stderr, "\n=== EXECUTION SUMMARY ===\n");
    fprintf(stderr, "Total queries processed: %zu\n", q);
    fprintf(stderr, "Total off-targets scored: %lu\n", static_cast<unsigned long>(total));
    fprintf(stderr, "Average off-targets per query: %.2f\n", q ? static_cast<double>(total) / q : 0.0);
    fprintf(stderr, "Loading time: %ld ms\n", static_cast<long>(loadMs));
    fprintf(stderr, "Processing time: %ld ms\n", static_cast<long>(procMs));
    fprintf(stderr, "Total execution time: %ld ms\n", static_cast<long>(loadMs + procMs));
    fprintf(stderr, "Processing rate: %.2f queries/sec\n", procMs > 0 ? q * 1000.0 / procMs : 0.0);
    fprintf(stderr, "Off-target scoring rate: %.2f off-targets/sec\n",
            procMs > 0 ? total * 1000.0 / procMs : 0.0);
    fprintf(stderr, "Distinct off-targets: %llu\n", static_cast<unsigned long long>(distinct));
    fprintf(stderr, "========================\n\n");

    std::string outPath;
    if (argc > 6) {
        outPath = argv[6];
    } else {
        outPath = std::filesystem::path(argv[2]).stem().string() + "_results.txt";
    }
    FILE *out = fopen(outPath.c_str(), "w");
    if (out)
        fprintf(out, "# Query\tMIT_Score\tCFD_Score\tOfftarget_Count\n");
    else
        fprintf(stderr, "Error: Could not open output file: %s\n", outPath.c_str());
    for (size_t i = 0; i < q; ++i) {
        const std::string seq = ISSLScoringModule::signatureToSequence(sigs[i], 20);
        printf("%s\t", seq.c_str());
        if (calcMit) printf("%f\t", mitv[i]); else printf("-1\t");
        if (calcCfd) printf("%f\n", cfdv[i]); else printf("-1\n");
        if (out) {
            fprintf(out, "%s\t", seq.c_str());
            if (calcMit) fprintf(out, "%f\t", mitv[i]); else fprintf(out, "-1\t");
            if (calcCfd) fprintf(out, "%f\t", cfdv[i]); else fprintf(out, "-1\t");
            fprintf(out, "%lu\n", static_cast<unsigned long>(cnt[i]));
        }
    }
    printf("Output time: %g seconds\n", secs(t_out, clk::now()));
    printf("Total execution time: %g seconds\n", secs(t_total, clk::now()));
    if (out) {
        fclose(out);
        fprintf(stderr, "Detailed results saved to: %s\n", outPath.c_str());
    }
    issl_index_close(ix);
    return 0;
}
