// Count-matrix writer in the 10x Genomics v2 directory layout that the
// reference's real-data comparison reads (experiments/compare_real_scrnaseq.jl:
// 62-96): matrix.mtx (MatrixMarket "coordinate integer general", genes x
// cells, 1-based indices, one nonzero per line), genes.tsv (id <TAB> name;
// the reader takes column 2 as the gene name) and barcodes.tsv.
//
// The counts arrive as the library's cells x G column-major int32 matrix, so
// entries are written gene by gene, cells ascending inside each gene: the
// nonzero count is taken in a first pass and the text is produced through one
// large buffer with a hand-rolled integer formatter.
#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "runtime_internal.h"

using namespace scrna;

namespace {

class Out {
  public:
    explicit Out(FILE* f) : f_(f) { buf_.reserve(1 << 22); }
    ~Out() { flush(); }
    void put(const char* s, size_t n) {
        buf_.append(s, n);
        if (buf_.size() >= (1u << 22)) flush();
    }
    void put(const std::string& s) { put(s.data(), s.size()); }
    void put(char c) { put(&c, 1); }
    void num(uint64_t v) {
        char tmp[24];
        int n = 0;
        do {
            tmp[n++] = (char)('0' + v % 10);
            v /= 10;
        } while (v);
        char out[24];
        for (int i = 0; i < n; ++i) out[i] = tmp[n - 1 - i];
        put(out, (size_t)n);
    }
    bool ok() const { return ok_; }
    void flush() {
        if (!buf_.empty() && std::fwrite(buf_.data(), 1, buf_.size(), f_) != buf_.size()) ok_ = false;
        buf_.clear();
    }

  private:
    FILE* f_;
    std::string buf_;
    bool ok_ = true;
};

// Deterministic 16-nt barcode of a cell (base-4 digits of the index, A/C/G/T),
// with the "-1" GEM-well suffix of 10x barcodes.
std::string barcode(const char* prefix, int64_t c) {
    static const char nt[4] = {'A', 'C', 'G', 'T'};
    std::string s = prefix ? prefix : "";
    char b[17];
    uint64_t v = (uint64_t)c;
    for (int i = 15; i >= 0; --i) {
        b[i] = nt[v & 3u];
        v >>= 2;
    }
    b[16] = '\0';
    s += b;
    s += "-1";
    return s;
}

int open_out(const std::string& path, FILE** f) {
    *f = std::fopen(path.c_str(), "wb");
    if (!*f) return set_error(SCRNA_ERR_INVALID, "cannot open %s: %s", path.c_str(), std::strerror(errno));
    return SCRNA_OK;
}

}  // namespace

int scrna_write_10x(const char* dir, const int32_t* counts, int64_t cells, int32_t G, int64_t ld,
                    const char* const* gene_names, const char* barcode_prefix, int64_t* nnz_out) {
    if (!dir || !counts) return set_error(SCRNA_ERR_INVALID, "null argument");
    if (cells < 1 || G < 1 || ld < cells) return set_error(SCRNA_ERR_INVALID, "need cells >= 1, G >= 1, ld >= cells");
    if (::mkdir(dir, 0755) != 0 && errno != EEXIST)
        return set_error(SCRNA_ERR_INVALID, "cannot create %s: %s", dir, std::strerror(errno));
    int64_t nnz = 0;
    for (int g = 0; g < G; ++g)
        for (int64_t c = 0; c < cells; ++c) {
            const int32_t v = counts[(size_t)g * ld + c];
            if (v < 0) return set_error(SCRNA_ERR_INVALID, "negative count at cell %lld gene %d", (long long)c, g);
            nnz += v != 0;
        }
    const std::string base(dir);
    FILE* f = nullptr;
    int rc;
    if ((rc = open_out(base + "/matrix.mtx", &f))) return rc;
    bool ok;
    {
        Out o(f);
        o.put(std::string("%%MatrixMarket matrix coordinate integer general\n"));
        o.put(std::string("% synthetic counts written by libscrna (genes x cells)\n"));
        o.num((uint64_t)G);
        o.put(' ');
        o.num((uint64_t)cells);
        o.put(' ');
        o.num((uint64_t)nnz);
        o.put('\n');
        for (int g = 0; g < G; ++g)
            for (int64_t c = 0; c < cells; ++c) {
                const int32_t v = counts[(size_t)g * ld + c];
                if (v == 0) continue;
                o.num((uint64_t)g + 1);
                o.put(' ');
                o.num((uint64_t)c + 1);
                o.put(' ');
                o.num((uint64_t)v);
                o.put('\n');
            }
        o.flush();
        ok = o.ok();
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return set_error(SCRNA_ERR_INVALID, "write failed: %s/matrix.mtx", dir);
    if ((rc = open_out(base + "/genes.tsv", &f))) return rc;
    {
        Out o(f);
        for (int g = 0; g < G; ++g) {
            char id[32];
            std::snprintf(id, sizeof id, "SYNG%06d", g + 1);
            o.put(std::string(id));
            o.put('\t');
            if (gene_names && gene_names[g])
                o.put(std::string(gene_names[g]));
            else
                o.put(std::string("gene") + std::to_string(g + 1));
            o.put('\n');
        }
        o.flush();
        ok = o.ok();
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return set_error(SCRNA_ERR_INVALID, "write failed: %s/genes.tsv", dir);
    if ((rc = open_out(base + "/barcodes.tsv", &f))) return rc;
    {
        Out o(f);
        for (int64_t c = 0; c < cells; ++c) {
            o.put(barcode(barcode_prefix, c));
            o.put('\n');
        }
        o.flush();
        ok = o.ok();
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return set_error(SCRNA_ERR_INVALID, "write failed: %s/barcodes.tsv", dir);
    if (nnz_out) *nnz_out = nnz;
    return SCRNA_OK;
}
