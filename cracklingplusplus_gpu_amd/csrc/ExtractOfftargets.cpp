// ExtractOfftargets.cpp -- off-target extraction CLI with the reference's
// arguments and output file (ExtractOfftargets/ExtractOfftargets.cpp:19-369):
//
//   ExtractOfftargets <output-file> {<input-file-1> ... <input-file-n> | <input-dir>}
//
// Inputs are FASTA or one sequence per line; the output file holds every
// site (20 bases followed by an NAG/NGG PAM, either strand), sorted, one per
// line, repeats kept -- the input of ISSLCreateIndex.  Scanning, sorting and
// formatting run on the GPU (issl_extract.h) instead of temporary files;
// stdout reports the same stages.  ISSL_DEVICE picks the GPU.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "issl_extract.h"

namespace {

[[noreturn]] void die(int rc, const char *what) {
    const char *m = issl_last_error();
    fprintf(stderr, "%s: %s (%d)\n", what, m ? m : "", rc);
    exit(1);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 3) {
        fprintf(stderr,
                "Usage: %s <output-file>  {<input-file-1> <input-file-2> ... <input-file-n> | "
                "<input-dir>}\n\n",
                argv[0]);
        exit(1);
    }
    const auto start = std::chrono::steady_clock::now();
    const char *dev = std::getenv("ISSL_DEVICE");
    issl_extractor *ex = nullptr;
    int rc = issl_extractor_create(dev ? atoi(dev) : 0, 0, &ex);
    if (rc != ISSL_OK) die(rc, "Error creating the extractor");

    printf("Spliting input(s)\nIdentifying off-targets\n");
    fflush(stdout);
    for (int i = 2; i < argc; ++i)
        if ((rc = issl_extractor_add_path(ex, argv[i])) != ISSL_OK) die(rc, argv[i]);
    uint64_t bases = 0, sites = 0, distinct = 0;
    issl_extractor_stats(ex, &bases, &sites);
    printf("Done\n");

    printf("Sorting off-targets\n");
    fflush(stdout);
    if ((rc = issl_extractor_finish(ex, &sites, &distinct)) != ISSL_OK) die(rc, "Sorting failed");
    printf("Done\n");

    printf("Writing %s\n", argv[1]);
    fflush(stdout);
    if ((rc = issl_extractor_write(ex, argv[1])) != ISSL_OK) die(rc, "Writing failed");
    printf("Done\n");
    issl_extractor_destroy(ex);

    const double secs =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
    printf("%llu bases, %llu off-targets (%llu distinct)\n", static_cast<unsigned long long>(bases),
           static_cast<unsigned long long>(sites), static_cast<unsigned long long>(distinct));
    const long long s = static_cast<long long>(secs);
    printf("Total run time %02lld %02lld:%02lld:%02lld (dd hh:mm:ss) or %lld seconds\n", s / 86400,
           (s / 3600) % 24, (s / 60) % 60, s % 60, s);
    return 0;
}
