// ISSLScoringModule.cpp -- the ISSL pipeline stage on the GPU, resident and
// memory-mapped.
//
// Mirrors lib/ISSLScoringModule.cpp: constructor method check (:13-50),
// run() (:52-466) with processGuide gating (:488-513), result fields and
// codes (:404-459), console messages; and lib/ISSLScoringModuleMMF.cpp, the
// same stage over a mapped index (:53-438).  Scoring goes through
// issl_score.
#include "ISSLScoringModule.hpp"
#include "ISSLScoringModuleMMF.hpp"

#include <cstdlib>
#include <iostream>
#include <locale>
#include <sstream>
#include <stdexcept>

#include "issl_gpu.h"

namespace {

std::string with_commas(uint64_t v) {
    std::string s = std::to_string(v);
    for (int i = static_cast<int>(s.size()) - 3; i > 0; i -= 3) s.insert(static_cast<size_t>(i), ",");
    return s;
}

void check(int rc, const char *what) {
    if (rc != ISSL_OK) {
        const char *m = issl_last_error();
        throw std::runtime_error(std::string(what) + ": " + (m ? m : ""));
    }
}

}  // namespace

ISSLScoringStage::ISSLScoringStage(const cracklingConfig &cfg)
    : config(cfg.offTarget),
      ISSLIndex(cfg.input.offtargetSites),
      toolIsSelected(cfg.offTarget.enabled),
      optimsationLevel(cfg.general.optimisation),
      consensusN(cfg.consensus.n) {
    switch (config.method) {
    case otScoreMethod::mitAndCfd:
    case otScoreMethod::mitOrCfd:
    case otScoreMethod::avgMitCfd:
    case otScoreMethod::mit:
    case otScoreMethod::cfd:
        break;
    default:
        throw std::runtime_error("Scoring method is not valid.");
    }
    if (const char *d = std::getenv("ISSL_DEVICE")) device = std::atoi(d);
}

ISSLScoringStage::~ISSLScoringStage() {
    if (index) issl_index_close(index);
}

void ISSLScoringStage::score(std::vector<guideResults> &candidateGuides, bool streamed) {
    if (!toolIsSelected) {
        std::cout << "Off-target scoring has been configured not to run. Skipping Off-target scoring"
                  << std::endl;
        return;
    }
    if (!index) {
        if (streamed) {
            // lib/ISSLScoringModuleMMF.cpp:61-130: the file is mapped, not
            // read; here its slices stream from the mapping through HBM
            // buffers (ISSL_STREAM_BYTES of them, default: what fits).
            std::cout << "Mapping ISSL Index." << std::endl;
            uint64_t budget = 0;
            if (const char *b = std::getenv("ISSL_STREAM_BYTES")) budget = std::strtoull(b, nullptr, 10);
            check(issl_index_open_streamed(ISSLIndex.string().c_str(), device, budget, &index),
                  "Error reading index");
        } else {
            std::cout << "Loading ISSL Index." << std::endl;
            check(issl_index_open(ISSLIndex.string().c_str(), device, &index), "Error reading index");
        }
        std::cout << "ISSL Index Loaded." << std::endl;
    }
    std::cout << "Beginning Off-target scoring." << std::endl;

    std::vector<size_t> picked;
    std::vector<uint64_t> sigs;
    picked.reserve(candidateGuides.size());
    for (size_t i = 0; i < candidateGuides.size(); ++i) {
        if (!processGuide(candidateGuides[i])) continue;
        picked.push_back(i);
        sigs.push_back(ISSLScoringModule::sequenceToSignature(candidateGuides[i].seq, 20));
    }
    const size_t n = picked.size();
    std::vector<double> mitv(n), cfdv(n);
    std::vector<char> pass(n);
    issl_params p{};
    p.max_dist = 4;
    p.method = static_cast<int32_t>(config.method);
    p.threshold = config.scoreThreshold;
    p.form = issl_form_from_env();
    p.precision = ISSL_PREC_FP64;
    issl_results r{};
    r.mit = mitv.data();
    r.cfd = cfdv.data();
    r.pass = pass.data();
    check(issl_score(index, sigs.data(), n, &p, &r), "ISSL scoring failed");

    uint64_t failed = 0;
    for (size_t k = 0; k < n; ++k) {
        guideResults &g = candidateGuides[picked[k]];
        g.mitOfftargetscore = mitv[k];
        g.cfdOfftargetscore = cfdv[k];
        g.passedOffTargetScore = pass[k] == '0' ? CODE_REJECTED : CODE_ACCEPTED;
        failed += pass[k] == '0';
    }
    std::cout << "\t" << with_commas(failed) << " of " << with_commas(n) << " failed here."
              << std::endl;
}

int issl_form_from_env() {
    const char *form = std::getenv("ISSL_FORM");
    const std::string f = form ? form : "";
    if (f == "batched") return ISSL_FORM_BATCHED;
    if (f == "plain") return ISSL_FORM_PLAIN;
    return ISSL_FORM_ORDERED;
}

ISSLScoringModule::ISSLScoringModule(const cracklingConfig &cfg) : ISSLScoringStage(cfg) {}

void ISSLScoringModule::run(std::vector<guideResults> &candidateGuides) {
    score(candidateGuides, false);
}

// lib/ISSLScoringModule.cpp:488-513 (the MMF module's :461-486 is the same)
bool ISSLScoringStage::processGuide(const guideResults &guide) const {
    if (optimsationLevel == optimisationLevel::ultralow) return true;
    if (!guide.isUnique) return false;
    if (optimsationLevel == optimisationLevel::medium || optimsationLevel == optimisationLevel::high) {
        if (guide.consensusCount < consensusN) return false;
        if (guide.passedBowtie2 == CODE_REJECTED) return false;
    }
    return true;
}

// lib/ISSLScoringModule.cpp:469-476 (characters other than C, G, T read as A)
uint64_t ISSLScoringModule::sequenceToSignature(const std::string &seq, uint64_t seqLen) {
    uint64_t sig = 0;
    for (uint64_t j = 0; j < seqLen && j < seq.size(); ++j) {
        const char c = seq[j];
        const uint64_t v = c == 'C' ? 1 : c == 'G' ? 2 : c == 'T' ? 3 : 0;
        sig |= v << (j * 2);
    }
    return sig;
}

std::string ISSLScoringModule::signatureToSequence(uint64_t sig, uint64_t seqLen) {
    std::string s(seqLen, ' ');
    for (uint64_t j = 0; j < seqLen; ++j) s[j] = "ACGT"[(sig >> (j * 2)) & 3];
    return s;
}

ISSLScoringModuleMMF::ISSLScoringModuleMMF(const cracklingConfig &cfg) : ISSLScoringStage(cfg) {}

void ISSLScoringModuleMMF::run(std::vector<guideResults> &candidateGuides) {
    score(candidateGuides, true);
}

uint64_t ISSLScoringModuleMMF::sequenceToSignature(const std::string &seq, uint64_t seqLen) {
    return ISSLScoringModule::sequenceToSignature(seq, seqLen);
}

std::string ISSLScoringModuleMMF::signatureToSequence(uint64_t sig, uint64_t seqLen) {
    return ISSLScoringModule::signatureToSequence(sig, seqLen);
}
