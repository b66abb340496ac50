// ISSLScoringModule.hpp -- drop-in replacement of the reference's ISSL
// pipeline stage (include/ISSLScoringModule.hpp:15-29): same constructor,
// run(std::vector<guideResults>&) and processGuide gating, scoring on the
// GPU through libissl_gpu.so.  ISSLScoringModuleMMF.hpp declares the
// memory-mapped variant on the same stage.
//
// Differences from the reference, all on the resource side:
//   * the index is opened once and stays resident in HBM across run() calls
//     (the reference re-reads the file inside every run,
//     lib/ISSLScoringModule.cpp:72-150);
//   * config.offTarget.threads does not apply (no OpenMP team);
//   * the ORDERED form scores by default: the reference's FP64 sums, counts
//     and exit point bit for bit; ISSL_FORM=batched in the environment
//     selects the BATCHED form (FP64, sums regrouped by tile) and
//     ISSL_FORM=plain the one-guide-per-workgroup PLAIN kernel (the same bits
//     as ORDERED).
#pragma once

#include <string>
#include <vector>

#include "crackling_types.hpp"

struct issl_index;

// Scoring form from the ISSL_FORM environment variable: "batched", "plain",
// else ORDERED (an ISSL_FORM_* value of issl_gpu.h).
int issl_form_from_env();

// The part both pipeline modules share: configuration, the processGuide
// gating (lib/ISSLScoringModule.cpp:488-513), the index handle and the
// scoring of one batch.  The reference classes derive privately from
// specificityModule; these derive privately from this stage.
class ISSLScoringStage {
protected:
    explicit ISSLScoringStage(const cracklingConfig &config);
    ~ISSLScoringStage();
    ISSLScoringStage(const ISSLScoringStage &) = delete;
    ISSLScoringStage &operator=(const ISSLScoringStage &) = delete;

    // Open the index on first use (resident, or streamed from a mapping of
    // the file) and score the batch's selected guides in place.
    void score(std::vector<guideResults> &candidateGuides, bool streamed);
    bool processGuide(const guideResults &guide) const;

    offTargetConfig config;
    std::filesystem::path ISSLIndex;
    bool toolIsSelected;
    optimisationLevel optimsationLevel;
    uint8_t consensusN;
    int device = 0;
    issl_index *index = nullptr;
};

class ISSLScoringModule : private ISSLScoringStage {
public:
    explicit ISSLScoringModule(const cracklingConfig &config);

    void run(std::vector<guideResults> &candidateGuides);

    // Signature helpers of lib/ISSLScoringModule.cpp:469-485.
    static uint64_t sequenceToSignature(const std::string &seq, uint64_t seqLen);
    static std::string signatureToSequence(uint64_t sig, uint64_t seqLen);
};
