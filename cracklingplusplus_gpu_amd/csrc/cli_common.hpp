// cli_common.hpp -- argument and query-file handling shared by the two
// standalone scorers (ISSLScoreOfftargets/ISSLScoreOfftargets.cpp:44-91,
// 205-288 and ISSLScoreOfftargetsMMF/ISSLScoreOfftargetsMMF.cpp:36-91,
// 205-240 read them the same way).
#pragma once

#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <string>
#include <vector>

#include "issl_gpu.h"

namespace issl_cli {

[[noreturn]] inline void die(int rc, const char *what) {
    const char *m = issl_last_error();
    fprintf(stderr, "%s: %s (%d)\n", what, m ? m : "", rc);
    exit(1);
}

// Score method argument: and | or | avg | mit | cfd.
inline void parse_method(const std::string &m, int *method, bool *calcMit, bool *calcCfd) {
    *calcMit = *calcCfd = true;
    if (m == "and") *method = ISSL_METHOD_MIT_AND_CFD;
    else if (m == "or") *method = ISSL_METHOD_MIT_OR_CFD;
    else if (m == "avg") *method = ISSL_METHOD_AVG_MIT_CFD;
    else if (m == "mit") { *method = ISSL_METHOD_MIT; *calcCfd = false; }
    else if (m == "cfd") { *method = ISSL_METHOD_CFD; *calcMit = false; }
    else {
        fprintf(stderr, "Invalid scoring method. Acceptable options are: 'and', 'or', 'avg', 'mit', 'cfd'");
        exit(1);
    }
}

// Query file of fixed-length lines (sequence + LF): encoded signatures.
inline std::vector<uint64_t> read_queries(const char *path, size_t lineLen) {
    std::error_code ec;
    const uintmax_t fileSize = std::filesystem::file_size(path, ec);
    if (ec) {
        fprintf(stderr, "Failed to read in query file.\n");
        exit(1);
    }
    if (fileSize % lineLen != 0) {
        fprintf(stderr, "Error: query file is not a multiple of the expected line length (%zu)\n", lineLen);
        fprintf(stderr, "The sequence length may be incorrect; alternatively, the line endings\n");
        fprintf(stderr, "may be something other than LF, or there may be junk at the end of the file.\n");
        exit(1);
    }
    const size_t q = fileSize / lineLen;
    std::vector<char> text(fileSize + 1);
    FILE *fp = fopen(path, "rb");
    if (!fp || (fileSize && fread(text.data(), fileSize, 1, fp) < 1)) {
        fprintf(stderr, "Failed to read in query file.\n");
        exit(1);
    }
    fclose(fp);
    if (lineLen < 21) {
        fprintf(stderr, "Error: the index holds sequences shorter than 20\n");
        exit(1);
    }
    std::vector<uint64_t> sigs(q);
    const int rc = issl_encode(text.data(), q, lineLen, sigs.data());
    if (rc != ISSL_OK) die(rc, "encode");
    return sigs;
}

}  // namespace issl_cli
