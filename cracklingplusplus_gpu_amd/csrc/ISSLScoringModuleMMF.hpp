// ISSLScoringModuleMMF.hpp -- drop-in replacement of the reference's
// memory-mapped ISSL stage (include/ISSLScoringModuleMMF.hpp:16-30), chosen
// by the Crackling pipeline when config.offTarget.memoryMapped is set
// (src/CracklingPlusPlus.cpp:73-80).  The .issl file is memory-mapped and
// its slices stream through HBM slice buffers (issl_index_open_streamed),
// so the index need not fit in device memory; results are the resident
// module's bit for bit.  ISSL_STREAM_BYTES sets the device budget of the
// slice buffers (default: as many as fit).
#pragma once

#include <string>
#include <vector>

#include "ISSLScoringModule.hpp"

class ISSLScoringModuleMMF : private ISSLScoringStage {
public:
    explicit ISSLScoringModuleMMF(const cracklingConfig &config);

    void run(std::vector<guideResults> &candidateGuides);

    // lib/ISSLScoringModuleMMF.cpp:442-459
    static uint64_t sequenceToSignature(const std::string &seq, uint64_t seqLen);
    static std::string signatureToSequence(uint64_t sig, uint64_t seqLen);
};
