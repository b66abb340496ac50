// ref_harness.cpp -- drives the reference's own ISSLScoringModule::run, built
// from the reference sources (lib/ISSLScoringModule.cpp, lib/otScorePenalties.cpp,
// lib/util.cpp, lib/specificityModule.cpp) by oracle/Makefile into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: used in this container to write golden vectors and
// to check the oracle; it is never part of the product.
//
//   issl_ref_harness tables <out.bin>
//       u64 n_mit, n_mit x (u64 key, f64 value) sorted by key,
//       320 f64 cfdPosPenalties, 16 f64 cfdPamPenalties
//   issl_ref_harness score <index.issl> <guides.txt> <method 0-4> <threshold> <threads> <out.bin>
//       per guide: f64 mitOfftargetscore, f64 cfdOfftargetscore, u8 passedOffTargetScore
//       guides.txt holds one guide (>= 20 characters) per line.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <utility>
#include <vector>

#include "ISSLScoringModule.hpp"
#include "otScorePenalties.hpp"
#include "util.hpp"

static int dump_tables(const char *out)
{
    std::vector<std::pair<uint64_t, double>> mit(precalculatedMITScores.begin(),
                                                 precalculatedMITScores.end());
    std::sort(mit.begin(), mit.end());
    FILE *f = fopen(out, "wb");
    if (!f) return 1;
    uint64_t n = mit.size();
    fwrite(&n, 8, 1, f);
    for (auto &kv : mit) {
        fwrite(&kv.first, 8, 1, f);
        fwrite(&kv.second, 8, 1, f);
    }
    fwrite(cfdPosPenalties, 8, 320, f);
    fwrite(cfdPamPenalties, 8, 16, f);
    fclose(f);
    return 0;
}

static int score(int argc, char **argv)
{
    if (argc != 8) return 2;
    std::vector<std::string> seqs;
    std::ifstream in(argv[3]);
    for (std::string line; std::getline(in, line);)
        if (line.size() >= 20) seqs.push_back(line);

    cracklingConfig cfg{};
    cfg.general.optimisation = optimisationLevel::ultralow;
    cfg.consensus.n = 0;
    cfg.input.offtargetSites = argv[2];
    cfg.offTarget.enabled = true;
    cfg.offTarget.memoryMapped = false;
    cfg.offTarget.method = static_cast<otScoreMethod>(atoi(argv[4]));
    cfg.offTarget.scoreThreshold = atof(argv[5]);
    cfg.offTarget.threads = static_cast<uint8_t>(atoi(argv[6]));

    std::vector<guideResults> guides(seqs.size());
    for (size_t i = 0; i < seqs.size(); ++i) {
        guides[i].seq = seqs[i];
        guides[i].isUnique = true;
    }
    ISSLScoringModule module(cfg);
    module.run(guides);

    FILE *f = fopen(argv[7], "wb");
    if (!f) return 1;
    for (auto &g : guides) {
        fwrite(&g.mitOfftargetscore, 8, 1, f);
        fwrite(&g.cfdOfftargetscore, 8, 1, f);
        uint8_t p = static_cast<uint8_t>(g.passedOffTargetScore);
        fwrite(&p, 1, 1, f);
    }
    fclose(f);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc >= 3 && std::string(argv[1]) == "tables") return dump_tables(argv[2]);
    if (argc >= 2 && std::string(argv[1]) == "score") return score(argc, argv);
    fprintf(stderr, "usage: %s tables <out> | score <issl> <guides> <method> <threshold> <threads> <out>\n", argv[0]);
    return 2;
}
