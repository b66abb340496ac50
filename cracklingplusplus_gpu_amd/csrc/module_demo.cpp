// module_demo.cpp -- drives the drop-in ISSLScoringModule (or, with
// memoryMapped set, ISSLScoringModuleMMF) the way the Crackling pipeline does
// (src/CracklingPlusPlus.cpp:38-39,73-80): build a cracklingConfig, fill
// guideResults, call run(), read the result fields.
//
//   issl_module_demo <index.issl> <guides.txt> <method> <threshold>
//                    <optimisation> <enabled 0|1> <consensus n> [mmf]
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
#include "ISSLScoringModuleMMF.hpp"

int main(int argc, char **argv) {
    if (argc != 8 && !(argc == 9 && std::string(argv[8]) == "mmf")) {
        fprintf(stderr, "usage: %s index guides method threshold optimisation enabled consensus [mmf]\n",
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
    cfg.offTarget.memoryMapped = argc == 9;

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
        // As the Crackling pipeline does: both modules exist, the flag picks one; a
        // second batch reuses the index opened by the first.
        ISSLScoringModule ISSLScoring(cfg);
        ISSLScoringModuleMMF ISSLScoringMMF(cfg);
        for (int batch = 0; batch < 2; ++batch) {
            if (cfg.offTarget.memoryMapped)
                ISSLScoringMMF.run(guides);
            else
                ISSLScoring.run(guides);
        }
    } catch (const std::exception &e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    for (const guideResults &g : guides)
        printf("RESULT\t%s\t%.17g\t%.17g\t%c\n", g.seq.c_str(), g.mitOfftargetscore,
               g.cfdOfftargetscore, g.passedOffTargetScore);
    return 0;
}
