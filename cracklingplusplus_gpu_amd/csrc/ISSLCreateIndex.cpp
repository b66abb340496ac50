// ISSLCreateIndex.cpp -- index builder CLI with the reference's argument list,
// checks and file layout (ISSLCreateIndex/ISSLCreateIndex.cpp:284-493):
//
//   ISSLCreateIndex <offtarget-sites> <slice-config> <sequence-length> <output-file>
//
// offtarget-sites: sorted sites, one sequence-length-mer per line (what
//   ExtractOfftargets writes), equal lines adjacent; every line ends with the
//   same line ending, detected after the first sequence as the reference does
//   (getLineEnding, :32-52): "\n", "\r\n", ...
// slice-config: one line per slice, '1' at position j puts j in the slice.
// Output: N | seqLength | S | sig[N] | sliceMask[S] | per slice: size[4^w]
//   then entries occ<<32 | id, keys ascending and ids ascending within a key.
//
// Counting and list building run on the host (OpenMP) through
// libissl_gpu.so; the file is memory-mapped, not read.  Differences: no
// temporary files (the reference writes <output>.tmp.1/.tmp.2), and a
// duplicate run at the end of the file is counted within the file (the
// reference compares past the last line, :398-401).  Sequence lengths above
// 21 are refused, as in the reference (:296); the scorer itself takes
// 20-position indexes only.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "issl_gpu.h"
#include "issl_index.h"

namespace {

struct Mapped {
    const char *p = nullptr;
    size_t len = 0;
    bool open(const char *path) {
        const int fd = ::open(path, O_RDONLY);
        if (fd < 0) return false;
        struct stat st {};
        if (fstat(fd, &st) != 0) {
            ::close(fd);
            return false;
        }
        len = static_cast<size_t>(st.st_size);
        if (len) {
            void *m = mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0);
            if (m == MAP_FAILED) {
                ::close(fd);
                return false;
            }
            p = static_cast<const char *>(m);
        }
        ::close(fd);
        return true;
    }
    ~Mapped() {
        if (p) munmap(const_cast<char *>(p), len);
    }
};

bool exists(const char *path) {
    struct stat st {};
    return ::stat(path, &st) == 0;
}

// The line ending of a file: the run of '\r' / '\n' characters right after
// its first seq_len characters (getLineEnding, ISSLCreateIndex.cpp:32-52).
std::string line_ending(const Mapped &m, uint64_t seq_len) {
    std::string e;
    for (size_t i = seq_len; i < m.len && (m.p[i] == '\r' || m.p[i] == '\n'); ++i) e.push_back(m.p[i]);
    return e;
}

std::string shown(const std::string &e) {
    std::string s;
    for (char c : e) s += c == '\r' ? "\\r" : c == '\n' ? "\\n" : std::string(1, c);
    return s;
}

// A file's size must be a whole number of lines (ISSLCreateIndex.cpp:312-337).
bool whole_lines(const char *what, const Mapped &m, uint64_t line_len, const std::string &ending) {
    if (line_len && m.len % line_len == 0) return true;
    fprintf(stderr, "fileSize: %zu\n", m.len);
    fprintf(stderr, "Error: %s file does is not a multiple of the expected line length (%llu)\n", what,
            static_cast<unsigned long long>(line_len));
    fprintf(stderr, "The sequence length may be incorrect; alternatively, the line endings\n");
    fprintf(stderr, "may be something other than %s, or there may be junk at the end of the file.\n",
            shown(ending).c_str());
    return false;
}

[[noreturn]] void die(const char *what) {
    fprintf(stderr, "%s: %s\n", what, issl_last_error());
    exit(1);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 5) {
        fprintf(stderr, "Usage: %s <offtarget-sites> <slice-config> <sequence-length> <output-file>\n\n",
                argv[0]);
        return 1;
    }
    const uint64_t seq_len = static_cast<uint64_t>(atoi(argv[3]));
    if (seq_len > 21) {
        fprintf(stderr, "Sequence length is greater than 21, which is the maximum supported currently\n\n");
        return 1;
    }
    if (!exists(argv[1])) {
        fprintf(stderr, "Could not find the specified offtarget sites file: %s\n", argv[1]);
        return 1;
    }
    Mapped sites;
    if (!sites.open(argv[1])) {
        fprintf(stderr, "Could not read the offtarget sites file: %s\n", argv[1]);
        return 1;
    }
    const std::string ot_end = line_ending(sites, seq_len);
    const uint64_t ot_line = seq_len + ot_end.size();
    if (!whole_lines("offtargetSites.txt", sites, ot_line, ot_end)) return 1;
    if (!exists(argv[2])) {
        fprintf(stderr, "Could not find the specified slice config file: %s\n", argv[2]);
        return 1;
    }
    Mapped slices;
    if (!slices.open(argv[2])) {
        fprintf(stderr, "Could not read the slice config file: %s\n", argv[2]);
        return 1;
    }
    const std::string sc_end = line_ending(slices, seq_len);
    if (!whole_lines("sliceconfig.txt", slices, seq_len + sc_end.size(), sc_end)) return 1;

    uint64_t n_slices = 0;
    if (issl_parse_masks(slices.p ? slices.p : "", slices.len, seq_len, nullptr, 0, &n_slices) != ISSL_OK)
        die("Error reading the slice config");
    std::vector<uint64_t> masks(n_slices);
    if (issl_parse_masks(slices.p ? slices.p : "", slices.len, seq_len, masks.data(), n_slices,
                         &n_slices) != ISSL_OK)
        die("Error reading the slice config");

    const uint64_t seq_count = sites.len / ot_line;
    printf("Number of sequences: %llu\n", static_cast<unsigned long long>(seq_count));
    printf("Counting occurrences...\n");
    uint64_t n = 0;
    if (issl_count_sorted_sites(sites.p ? sites.p : "", sites.len, seq_len, ot_line, nullptr, nullptr,
                                0, &n) != ISSL_OK)
        die("Error counting the offtarget sites");
    std::vector<uint64_t> sigs(n);
    std::vector<uint32_t> occ(n);
    if (issl_count_sorted_sites(sites.p ? sites.p : "", sites.len, seq_len, ot_line, sigs.data(),
                                occ.data(), n, &n) != ISSL_OK)
        die("Error counting the offtarget sites");
    printf("Finished!\n");

    printf("Constructing index...\n");
    const uint64_t n_lists = issl_list_count(masks.data(), n_slices, seq_len);
    if (n_slices && n_lists == 0) {
        fprintf(stderr, "Error: a slice is wider than 16 positions\n");
        return 1;
    }
    std::vector<uint64_t> sizes(n_lists), entries(n_slices * n);
    if (issl_build_lists(sigs.data(), occ.data(), n, seq_len, masks.data(), n_slices, sizes.data(),
                         entries.data(), 0) != ISSL_OK)
        die("Error building the slice lists");
    printf("Finished!\n");
    printf("Writing index to file...\n");
    if (issl_write_index(argv[4], n, seq_len, n_slices, sigs.data(), masks.data(), sizes.data(),
                         entries.data()) != ISSL_OK)
        die("Error writing the index");
    printf("Finished!\n");
    printf("Done\n");
    return 0;
}
