// formats.cpp -- the on-disk formats of dcsr_matrix_t (HLanc.hpp:103-392).
//
// "pp" (one row per line of `col:val` tokens), Matrix Market coordinate ("mm")
// and the raw binary CSR ("bin": int ncols, int nrows, int row_ptr[nrows+1],
// int nnz, int col[nnz], double val[nnz]).  Reading rules follow the reference:
// zero values are dropped, a symmetric matrix keeps its lower triangle (upper
// entries are dropped with a warning), entries keep file order within a row.
// Where the reference prints and exits, these return HL_ERR_IO / HL_ERR_FORMAT
// with the same message in hl_last_error().
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "hl_internal.hpp"

namespace hl {
namespace {

struct Rows {
    std::vector<std::vector<int>> c;
    std::vector<std::vector<double>> v;
};

int fill(hl_matrix *m, int nrows, int ncols, bool sym, const std::vector<int> &rp, const std::vector<int> &col,
         const std::vector<double> &val)
{
    m->nrows = nrows;
    m->ncols = ncols;
    m->nnz = static_cast<int>(val.size());
    m->sym = sym ? 1 : 0;
    m->row_ptr = static_cast<int *>(std::malloc(sizeof(int) * (rp.size() ? rp.size() : 1)));
    m->col = static_cast<int *>(std::malloc(sizeof(int) * (col.size() ? col.size() : 1)));
    m->val = static_cast<double *>(std::malloc(sizeof(double) * (val.size() ? val.size() : 1)));
    if (!m->row_ptr || !m->col || !m->val) return HL_ERR_OOM;
    std::copy(rp.begin(), rp.end(), m->row_ptr);
    std::copy(col.begin(), col.end(), m->col);
    std::copy(val.begin(), val.end(), m->val);
    return HL_OK;
}

// HLanc.hpp:119-169
int read_pp(const char *fn, bool sym, hl_matrix *out)
{
    FILE *fp = std::fopen(fn, "r");
    if (!fp) {
        set_error("Error opening file %s: %s", fn, std::strerror(errno));
        return HL_ERR_IO;
    }
    std::vector<int> rp{0}, col;
    std::vector<double> val;
    int ncols = 0;
    for (;;) {
        int ch;
        while ((ch = std::fgetc(fp)) != EOF && (ch == ' ' || ch == '\t')) {
        }
        if (ch == '\n') {
            rp.push_back(static_cast<int>(val.size()));
            continue;
        }
        if (ch == EOF) break;
        std::ungetc(ch, fp);
        int idx;
        double v;
        if (std::fscanf(fp, "%d:%lf", &idx, &v) != 2) {
            std::fclose(fp);
            set_error("Error reading file %s at line %d.", fn, static_cast<int>(rp.size()));
            return HL_ERR_FORMAT;
        }
        if (v == 0.0) continue;
        const int row = static_cast<int>(rp.size()) - 1;
        if (sym && idx > row) {
            std::fprintf(stderr, "Ignore element [%d, %d] = %lg\n", row, idx, v);
        } else {
            col.push_back(idx);
            val.push_back(v);
        }
        if (ncols < idx) ncols = idx;
        if (sym && ncols < row) ncols = row;
    }
    std::fclose(fp);
    const int nrows = static_cast<int>(rp.size()) - 1;
    if (static_cast<int>(val.size()) != rp.back()) { // last line had no newline
        std::fprintf(stderr, "Ignore %d element(s) of an unterminated last line in %s\n",
                     static_cast<int>(val.size()) - rp.back(), fn);
        col.resize(rp.back());
        val.resize(rp.back());
    }
    return fill(out, nrows, ncols + 1, sym, rp, col, val);
}

// HLanc.hpp:171-247
int read_mm(const char *fn, hl_matrix *out)
{
    FILE *fp = std::fopen(fn, "r");
    if (!fp) {
        set_error("Error opening file %s: %s", fn, std::strerror(errno));
        return HL_ERR_IO;
    }
    char *line = nullptr;
    size_t len = 0;
    if (getline(&line, &len, fp) == -1) {
        std::fclose(fp);
        std::free(line);
        set_error("Error reading banner in file %s", fn);
        return HL_ERR_FORMAT;
    }
    bool sym = false;
    if (std::strstr(line, "%%MatrixMarket matrix coordinate") != line) {
        set_error("Unknown banner format: %s", line);
        std::free(line);
        std::fclose(fp);
        return HL_ERR_FORMAT;
    }
    if (std::strstr(line, "symmetric")) {
        sym = true;
    } else if (!std::strstr(line, "general")) {
        set_error("Unknown banner format: %s", line);
        std::free(line);
        std::fclose(fp);
        return HL_ERR_FORMAT;
    }
    std::free(line);
    int ch;
    while ((ch = std::fgetc(fp)) == '%') {
        while ((ch = std::fgetc(fp)) != '\n' && ch != EOF) {
        }
    }
    if (ch != EOF) std::ungetc(ch, fp);
    int rows, cols, entries;
    if (std::fscanf(fp, "%d%d%d", &rows, &cols, &entries) != 3 || rows < 0 || cols < 0 || entries < 0) {
        std::fclose(fp);
        set_error("Error reading rows, columns and entries");
        return HL_ERR_FORMAT;
    }
    Rows R;
    R.c.resize(rows);
    R.v.resize(rows);
    while (entries--) {
        int r, c;
        double v;
        if (std::fscanf(fp, "%d%d%lf", &r, &c, &v) != 3 || r < 1 || r > rows || c < 1 || c > cols) {
            std::fclose(fp);
            set_error("Error reading element");
            return HL_ERR_FORMAT;
        }
        if (v == 0.0) continue;
        R.c[r - 1].push_back(c - 1);
        R.v[r - 1].push_back(v);
    }
    std::fclose(fp);
    std::vector<int> rp{0}, col;
    std::vector<double> val;
    for (int r = 0; r < rows; ++r) {
        for (size_t i = 0; i < R.c[r].size(); ++i) {
            if (sym && R.c[r][i] > r) {
                std::fprintf(stderr, "Ignore element [%d, %d] = %lg\n", r, R.c[r][i], R.v[r][i]);
                continue;
            }
            col.push_back(R.c[r][i]);
            val.push_back(R.v[r][i]);
        }
        rp.push_back(static_cast<int>(val.size()));
    }
    return fill(out, rows, cols, sym, rp, col, val);
}

// HLanc.hpp:249-285
int read_bin(const char *fn, hl_matrix *out)
{
    FILE *fp = std::fopen(fn, "rb");
    if (!fp) {
        set_error("Error opening file %s: %s", fn, std::strerror(errno));
        return HL_ERR_IO;
    }
    int ncols = 0, nrows = 0, nnz = 0;
    auto fail = [&](const char *what) {
        std::fclose(fp);
        set_error("Error reading %s.", what);
        return HL_ERR_FORMAT;
    };
    if (std::fread(&ncols, sizeof(int), 1, fp) != 1) return fail("num_columns");
    if (std::fread(&nrows, sizeof(int), 1, fp) != 1 || nrows < 0) return fail("num_rows");
    std::vector<int> rp(static_cast<size_t>(nrows) + 1);
    if (std::fread(rp.data(), sizeof(int), rp.size(), fp) != rp.size()) return fail("csr_row_ptr[]");
    if (std::fread(&nnz, sizeof(int), 1, fp) != 1 || nnz < 0) return fail("nnz");
    std::vector<int> col(nnz);
    std::vector<double> val(nnz);
    if (std::fread(col.data(), sizeof(int), nnz, fp) != static_cast<size_t>(nnz)) return fail("csr_column_index[]");
    if (std::fread(val.data(), sizeof(double), nnz, fp) != static_cast<size_t>(nnz)) return fail("csr_value[]");
    std::fclose(fp);
    return fill(out, nrows, ncols, false, rp, col, val);
}

} // namespace
} // namespace hl

using namespace hl;

extern "C" {

int hl_matrix_read(const char *format, const char *filename, int is_symmetric, hl_matrix *out)
try {
    if (!format || !filename || !out) return HL_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    int rc;
    if (!std::strcmp(format, "pp"))
        rc = read_pp(filename, is_symmetric != 0, out);
    else if (!std::strcmp(format, "mm"))
        rc = read_mm(filename, out);
    else if (!std::strcmp(format, "bin"))
        rc = read_bin(filename, out);
    else {
        set_error("Unsupported format: %s", format);
        return HL_ERR_ARG;
    }
    if (rc) hl_matrix_free(out);
    return rc;
} catch (...) {
    return exception_code();
}

void hl_matrix_free(hl_matrix *m)
{
    if (!m) return;
    std::free(m->row_ptr);
    std::free(m->col);
    std::free(m->val);
    std::memset(m, 0, sizeof(*m));
}

int hl_matrix_write(const char *format, const char *filename, const hl_matrix *m)
try {
    if (!format || !filename || !m) return HL_ERR_ARG;
    const int nr = m->nrows;
    if (!std::strcmp(format, "bin")) {
        // save_bin expands symmetric storage (HLanc.hpp:339-364)
        std::vector<int> rp(static_cast<size_t>(nr) + 1, 0), col;
        std::vector<double> val;
        if (m->sym) {
            std::vector<long long> cnt(static_cast<size_t>(nr) + 1, 0);
            for (int r = 0; r < nr; ++r)
                for (int q = m->row_ptr[r]; q < m->row_ptr[r + 1]; ++q) {
                    cnt[r + 1]++;
                    if (m->col[q] != r) cnt[m->col[q] + 1]++;
                }
            for (int r = 0; r < nr; ++r) cnt[r + 1] += cnt[r];
            for (int r = 0; r <= nr; ++r) rp[r] = static_cast<int>(cnt[r]);
            col.resize(cnt[nr]);
            val.resize(cnt[nr]);
            std::vector<long long> cur(cnt.begin(), cnt.end() - 1);
            for (int r = 0; r < nr; ++r)
                for (int q = m->row_ptr[r]; q < m->row_ptr[r + 1]; ++q) {
                    const int c = m->col[q];
                    col[cur[r]] = c;
                    val[cur[r]++] = m->val[q];
                    if (c != r) {
                        col[cur[c]] = r;
                        val[cur[c]++] = m->val[q];
                    }
                }
        } else {
            rp.assign(m->row_ptr, m->row_ptr + nr + 1);
            col.assign(m->col, m->col + m->nnz);
            val.assign(m->val, m->val + m->nnz);
        }
        FILE *fp = std::fopen(filename, "wb");
        if (!fp) {
            set_error("Error opening file %s: %s", filename, std::strerror(errno));
            return HL_ERR_IO;
        }
        const int nnz = static_cast<int>(col.size());
        bool ok = std::fwrite(&m->ncols, sizeof(int), 1, fp) == 1 && std::fwrite(&nr, sizeof(int), 1, fp) == 1 &&
                  std::fwrite(rp.data(), sizeof(int), rp.size(), fp) == rp.size() &&
                  std::fwrite(&nnz, sizeof(int), 1, fp) == 1 &&
                  std::fwrite(col.data(), sizeof(int), nnz, fp) == static_cast<size_t>(nnz) &&
                  std::fwrite(val.data(), sizeof(double), nnz, fp) == static_cast<size_t>(nnz);
        std::fclose(fp);
        if (!ok) {
            set_error("Error writing file %s", filename);
            return HL_ERR_IO;
        }
        return HL_OK;
    }
    const bool pp = !std::strcmp(format, "pp");
    if (!pp && std::strcmp(format, "mm")) {
        set_error("Unsupported format: %s", format);
        return HL_ERR_ARG;
    }
    FILE *fp = std::fopen(filename, "w");
    if (!fp) {
        set_error("Error opening file %s: %s", filename, std::strerror(errno));
        return HL_ERR_IO;
    }
    bool ok = true;
    if (pp) { // save_pp (HLanc.hpp:287-303): rows without entries write nothing
        for (int r = 0; r < nr && ok; ++r)
            for (int i = m->row_ptr[r]; i < m->row_ptr[r + 1] && ok; ++i)
                ok = std::fprintf(fp, "%d:%lg%c", m->col[i], m->val[i], i + 1 == m->row_ptr[r + 1] ? '\n' : ' ') >= 0;
    } else { // save_mm (HLanc.hpp:305-328)
        ok = std::fprintf(fp, "%%%%MatrixMarket matrix coordinate real %s\n", m->sym ? "symmetric" : "general") >= 0 &&
             std::fprintf(fp, "%d %d %d\n", nr, m->ncols, m->nnz) >= 0;
        for (int r = 0; r < nr && ok; ++r)
            for (int i = m->row_ptr[r]; i < m->row_ptr[r + 1] && ok; ++i)
                ok = std::fprintf(fp, "%d %d %lg\n", r + 1, m->col[i] + 1, m->val[i]) >= 0;
    }
    std::fclose(fp);
    if (!ok) {
        set_error("Error writing file %s", filename);
        return HL_ERR_IO;
    }
    return HL_OK;
} catch (...) {
    return exception_code();
}

} // extern "C"
