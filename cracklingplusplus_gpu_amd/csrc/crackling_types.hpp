// crackling_types.hpp -- the Crackling pipeline types the ISSL stage reads and
// writes, with the member names and meanings of the reference's
// include/util.hpp:20-172 (enums otScoreMethod/optimisationLevel, result codes
// of lib/util.cpp:21-25, cracklingConfig and guideResults).
//
// Inside the Crackling tree, build with -DISSL_CRACKLING_UTIL_HPP='"util.hpp"'
// so the drop-in module uses the pipeline's own definitions instead.
#pragma once

#ifdef ISSL_CRACKLING_UTIL_HPP
#include ISSL_CRACKLING_UTIL_HPP
#else

#include <cfloat>
#include <climits>
#include <cstdint>
#include <filesystem>
#include <string>
#include <vector>

enum otScoreMethod { mit = 0, cfd = 1, mitAndCfd = 2, mitOrCfd = 3, avgMitCfd = 4 };
enum optimisationLevel { ultralow = 0, low = 1, medium = 2, high = 3 };

inline constexpr char CODE_ACCEPTED = '1';
inline constexpr char CODE_REJECTED = '0';
inline constexpr char CODE_UNTESTED = '?';
inline constexpr char CODE_AMBIGUOUS = '-';
inline constexpr char CODE_ERROR = '!';

struct generalConfig {
    std::string name;
    optimisationLevel optimisation;
};

struct consensusConfig {
    uint8_t n;
    uint8_t toolCount;
    bool mm10db;
    bool sgrnascorer2;
    bool chopchop;
};

struct inputConfig {
    std::filesystem::path exonSequences;
    std::filesystem::path offtargetSites;
    std::filesystem::path bowtie2Index;
    std::vector<std::filesystem::path> filesToProcess;
    uint64_t batchLen;
};

struct offTargetConfig {
    bool enabled;
    bool memoryMapped;
    otScoreMethod method;
    uint8_t threads;
    uint64_t pageLen;
    double scoreThreshold;
};

struct cracklingConfig {
    generalConfig general;
    consensusConfig consensus;
    inputConfig input;
    offTargetConfig offTarget;
};

struct guideResults {
    std::string seq = std::string(1, CODE_AMBIGUOUS);
    std::string header = std::string(1, CODE_AMBIGUOUS);
    uint64_t start = ULLONG_MAX;
    uint64_t end = ULLONG_MAX;
    char strand = CODE_AMBIGUOUS;
    bool isUnique = false;
    char passedG20 = CODE_UNTESTED;
    char passedAvoidLeadingT = CODE_UNTESTED;
    char passedATPercent = CODE_UNTESTED;
    char passedTTTT = CODE_UNTESTED;
    char passedSecondaryStructure = CODE_UNTESTED;
    char acceptedByMm10db = CODE_UNTESTED;
    char acceptedBySgRnaScorer2 = CODE_UNTESTED;
    int8_t consensusCount = 0;
    char passedBowtie2 = CODE_UNTESTED;
    char passedOffTargetScore = CODE_UNTESTED;
    double AT = -DBL_MAX;
    std::string ssL1 = std::string(1, CODE_UNTESTED);
    std::string ssStructure = std::string(1, CODE_UNTESTED);
    double ssEnergy = -DBL_MAX;
    double sgrnascorer2score = -DBL_MAX;
    std::string bowtie2Chr = std::string(1, CODE_UNTESTED);
    uint64_t bowtie2Start = ULLONG_MAX;
    uint64_t bowtie2End = ULLONG_MAX;
    double mitOfftargetscore = -DBL_MAX;
    double cfdOfftargetscore = -DBL_MAX;
};

#endif
