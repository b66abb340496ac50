// ISSLScoreOfftargetsMMF.cpp -- the memory-mapped standalone scorer with the
// reference's arguments and output (ISSLScoreOfftargetsMMF/ISSLScoreOfftargetsMMF.cpp:34-457):
//
//   ISSLScoreOfftargetsMMF [issltable] [query file] [max distance]
//                          [score-threshold] [score-method]
//
// stdout: "SEQ\tMIT\tCFD" per query (%f, or -1 when not computed), nothing
// else (the reference prints only these lines, :440-455).
//
// The index file is memory-mapped and its slices stream through HBM slice
// buffers (issl_index_open_streamed), so it need not fit in device memory;
// ISSL_STREAM_BYTES sets their budget (default: as many as fit).  Results
// are those of ISSLScoreOfftargets bit for bit.  Five arguments are required
// (the reference checks argc < 4 but reads argv[4] and argv[5]); the
// distance limit is 4 whatever the third argument says (:369).
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "ISSLScoringModule.hpp"
#include "cli_common.hpp"
#include "issl_gpu.h"

int main(int argc, char **argv) {
    if (argc < 6) {
        fprintf(stderr, "Usage: %s [issltable] [query file] [max distance] [score-threshold] [score-method]\n",
                argv[0]);
        exit(1);
    }
    const double threshold = atof(argv[4]);
    int method;
    bool calcMit, calcCfd;
    issl_cli::parse_method(argv[5], &method, &calcMit, &calcCfd);

    issl_index *ix = nullptr;
    const char *dev = std::getenv("ISSL_DEVICE");
    const char *budget = std::getenv("ISSL_STREAM_BYTES");
    int rc = issl_index_open_streamed(argv[1], dev ? atoi(dev) : 0,
                                      budget ? std::strtoull(budget, nullptr, 10) : 0, &ix);
    if (rc != ISSL_OK) issl_cli::die(rc, "Error reading index");
    issl_index_info info{};
    issl_index_get_info(ix, &info);
    const std::vector<uint64_t> sigs = issl_cli::read_queries(argv[2], info.seq_len + 1);
    const size_t q = sigs.size();

    std::vector<double> mitv(q), cfdv(q);
    issl_params p{};
    p.max_dist = 4;
    p.method = method;
    p.threshold = threshold;
    p.form = ISSL_FORM_PLAIN;
    p.precision = ISSL_PREC_FP64;
    issl_results r{};
    r.mit = mitv.data();
    r.cfd = cfdv.data();
    rc = issl_score(ix, sigs.data(), q, &p, &r);
    if (rc != ISSL_OK) issl_cli::die(rc, "ISSL scoring failed");
    for (size_t i = 0; i < q; ++i) {
        const std::string seq = ISSLScoringModule::signatureToSequence(sigs[i], 20);
        printf("%s\t", seq.c_str());
        if (calcMit) printf("%f\t", mitv[i]); else printf("-1\t");
        if (calcCfd) printf("%f\n", cfdv[i]); else printf("-1\n");
    }
    issl_index_close(ix);
    return 0;
}
