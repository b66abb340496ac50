// module_demo.cpp -- drives the drop-in ISSLScoringModule the way the
// Crackling driver does (src/CracklingPlusPlus.cpp:38,74-81): build a
// cracklingConfig, fill guideResults, call run(), read the result fields.
//
//   issl_module_demo <index.issl> <guides.txt> <method> <threshold>
//                    <optimisation> <enabled 0|1> <consensus n>
//
// guides.txt: one guide per line: "SEQ isUnique(0|1) consensusCount passedBowtie2(char)".
// Output (stdout): "SEQ\t<mit %.17g>\t<cfd %.17g>\t<passedOffTargetScore>" per guide.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "ISSLScoringModule.hpp"

int main(int argc, char **argv) {
    if (argc != 8) {
        fprintf(stderr, "usage: %s index guides method threshold optimisation enabled consensus\n",
                argv[0]);
        return 2;
    }
    const std::map<std::string, otScoreMethod> methods = {
        {"mit", otScoreMethod::mit},           {"cfd", otScoreMethod::cfd},
        {"mitAndCfd", otScoreMethod::mitAndCfd}, {"mitOrCfd", otScoreMethod::mitOrCfd},
        {"avgMitCfd", otScoreMethod::avgMitCfd}};
    const std::map<std::string, optimisationLevel> levels = {
        {"ultralow", optimisationLevel::ultralow}, {"low", optimisationLevel::low},
        {"medium", optimisationLevel::medium},     {"high", optimisationLevel::high}};
    cracklingConfig cfg{};
    cfg.input.offtargetSites = argv[1];
    auto mi = methods.find(argv[3]);
    cfg.offTarget.method = mi == methods.end() ? static_cast<otScoreMethod>(99) : mi->second;
    cfg.offTarget.scoreThreshold = atof(argv[4]);
    cfg.general.optimisation = levels.at(argv[5]);
    cfg.offTarget.enabled = atoi(argv[6]) != 0;
    cfg.consensus.n = static_cast<uint8_t>(atoi(argv[7]));

    std::vector<guideResults> guides;
    std::ifstream in(argv[2]);
    for (std::string line; std::getline(in, line);) {
        std::istringstream ls(line);
        guideResults g;
        int uniq = 1, cons = 0;
        std::string bt = "?";
        ls >> g.seq >> uniq >> cons >> bt;
        g.isUnique = uniq != 0;
        g.consensusCount = static_cast<int8_t>(cons);
        g.passedBowtie2 = bt.empty() ? '?' : bt[0];
        guides.push_back(g);
    }
    try {
        ISSLScoringModule module(cfg);
        module.run(guides);
        module.run(guides);  // a second batch reuses the resident index
    } catch (const std::exception &e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    for (const guideResults &g : guides)
        printf("RESULT\t%s\t%.17g\t%.17g\t%c\n", g.seq.c_str(), g.mitOfftargetscore,
               g.cfdOfftargetscore, g.passedOffTargetScore);
    return 0;
}
