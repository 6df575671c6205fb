// qs_io.hip -- device states to and from files in qpp's own formats
// (include/qpp/input_output.hpp), so states travel between qpp and the
// engine.
//
//   text   (qpp::save / qpp::load<cmat>, input_output.hpp:177-211, 233-262):
//          "rows cols\n", then for every row its entries "(re,im)" separated
//          by one space and ended by '\n'; each real printed with
//          max_digits10 = 17 significant digits (internal/util.hpp:440-446,
//          std::stringstream << std::setprecision(17), the same output as
//          "%.17g"), which round-trips every double exactly.
//   binary (qpp::obsolete::save / load, input_output.hpp:318-345, 357-390):
//          the 19 bytes "TYPE::Eigen::Matrix", rows and cols as 8-byte
//          size_t, then the column-major complex<double> data.
//
// A ket is a D x 1 matrix, a density matrix D x D (column-major, as the
// state is stored).  Kets stream through a bounded host buffer in chunks; a
// density matrix is written row by row, so its text form stages the whole
// matrix on the host.  complex64 states are written as complex<double>
// (qpp's cplx) and converted back on load.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "qs_internal.hpp"

namespace {

using qs::fail;
using cd = std::complex<double>;

constexpr uint64_t kChunk = uint64_t{1} << 22;  // elements per host transfer
const char kHeader[] = "TYPE::Eigen::Matrix";
constexpr size_t kHeaderLen = sizeof(kHeader) - 1;

struct FileCloser {
    void operator()(std::FILE* f) const {
        if (f) std::fclose(f);
    }
};
using File = std::unique_ptr<std::FILE, FileCloser>;

// elements [e0, e0 + n) of the state as complex<double>
int fetch(const qs_state* st, uint64_t e0, uint64_t n, cd* out, std::vector<float>& tmp,
          const char* W) {
    const uint64_t es = st->elem_bytes();
    const char* src = static_cast<const char*>(st->d) + e0 * es;
    if (st->dtype == QS_C128) {
        QS_HIP(hipMemcpyAsync(out, src, n * es, hipMemcpyDeviceToHost, st->stream), W);
        QS_HIP(hipStreamSynchronize(st->stream), W);
        return QS_OK;
    }
    tmp.resize(2 * n);
    QS_HIP(hipMemcpyAsync(tmp.data(), src, n * es, hipMemcpyDeviceToHost, st->stream), W);
    QS_HIP(hipStreamSynchronize(st->stream), W);
    for (uint64_t i = 0; i < n; ++i) out[i] = cd(tmp[2 * i], tmp[2 * i + 1]);
    return QS_OK;
}

// elements [e0, e0 + n) of the state from complex<double>
int store(qs_state* st, uint64_t e0, uint64_t n, const cd* in, std::vector<float>& tmp,
          const char* W) {
    const uint64_t es = st->elem_bytes();
    char* dst = static_cast<char*>(st->d) + e0 * es;
    if (st->dtype == QS_C128) {
        QS_HIP(hipMemcpyAsync(dst, in, n * es, hipMemcpyHostToDevice, st->stream), W);
        QS_HIP(hipStreamSynchronize(st->stream), W);
        return QS_OK;
    }
    tmp.resize(2 * n);
    for (uint64_t i = 0; i < n; ++i) {
        tmp[2 * i] = static_cast<float>(in[i].real());
        tmp[2 * i + 1] = static_cast<float>(in[i].imag());
    }
    QS_HIP(hipMemcpyAsync(dst, tmp.data(), n * es, hipMemcpyHostToDevice, st->stream), W);
    QS_HIP(hipStreamSynchronize(st->stream), W);
    return QS_OK;
}

// "(re,im)" with max_digits10 significant digits, as qpp::save prints it
size_t put_entry(char* buf, size_t cap, const cd& z) {
    const int k = std::snprintf(buf, cap, "(%.17g,%.17g)", z.real(), z.imag());
    return k > 0 ? static_cast<size_t>(k) : 0;
}

int save_text(const qs_state* st, std::FILE* f, uint64_t rows, uint64_t cols, const char* W) {
    std::vector<float> tmp;
    if (std::fprintf(f, "%llu %llu\n", static_cast<unsigned long long>(rows),
                     static_cast<unsigned long long>(cols)) < 0)
        return qs::fail_runtime(QS_ERR_IO, "qpp::save(): Error writing output stream!");
    std::string line;
    char ent[96];
    if (cols == 1) {  // ket: one entry per line, streamed in chunks
        std::vector<cd> h(std::min<uint64_t>(rows, kChunk));
        for (uint64_t e0 = 0; e0 < rows; e0 += h.size()) {
            const uint64_t n = std::min<uint64_t>(h.size(), rows - e0);
            if (int rc = fetch(st, e0, n, h.data(), tmp, W)) return rc;
            line.clear();
            for (uint64_t i = 0; i < n; ++i) {
                line.append(ent, put_entry(ent, sizeof ent, h[i]));
                line.push_back('\n');
            }
            if (std::fwrite(line.data(), 1, line.size(), f) != line.size())
                return qs::fail_runtime(QS_ERR_IO, "qpp::save(): Error writing output stream!");
        }
        return QS_OK;
    }
    // density matrix: rows of a column-major matrix, staged whole
    std::vector<cd> h(rows * cols);
    for (uint64_t e0 = 0; e0 < h.size(); e0 += kChunk) {
        const uint64_t n = std::min<uint64_t>(kChunk, h.size() - e0);
        if (int rc = fetch(st, e0, n, h.data() + e0, tmp, W)) return rc;
    }
    for (uint64_t i = 0; i < rows; ++i) {
        line.clear();
        for (uint64_t j = 0; j < cols; ++j) {
            if (j) line.push_back(' ');
            line.append(ent, put_entry(ent, sizeof ent, h[i + j * rows]));
        }
        line.push_back('\n');
        if (std::fwrite(line.data(), 1, line.size(), f) != line.size())
            return qs::fail_runtime(QS_ERR_IO, "qpp::save(): Error writing output stream!");
    }
    return QS_OK;
}

int save_binary(const qs_state* st, std::FILE* f, uint64_t rows, uint64_t cols, const char* W) {
    const char* E = "qpp::obsolete::save(): Error writing output stream!";
    std::vector<float> tmp;
    static_assert(sizeof(size_t) == 8, "qpp writes idx = std::size_t");
    const size_t r = rows, c = cols;
    if (std::fwrite(kHeader, 1, kHeaderLen, f) != kHeaderLen || std::fwrite(&r, 8, 1, f) != 1 ||
        std::fwrite(&c, 8, 1, f) != 1)
        return qs::fail_runtime(QS_ERR_IO, E);
    const uint64_t total = rows * cols;
    std::vector<cd> h(std::min<uint64_t>(total, kChunk));
    for (uint64_t e0 = 0; e0 < total; e0 += h.size()) {
        const uint64_t n = std::min<uint64_t>(h.size(), total - e0);
        if (int rc = fetch(st, e0, n, h.data(), tmp, W)) return rc;
        if (std::fwrite(h.data(), sizeof(cd), n, f) != n) return qs::fail_runtime(QS_ERR_IO, E);
    }
    return QS_OK;
}

int load_text(qs_state* st, std::FILE* f, uint64_t rows, uint64_t cols, const char* W) {
    const char* E = "qpp::load(): Input stream is corrupted!";
    std::vector<float> tmp;
    auto entry = [&](cd& z) {
        // is >> skip >> re >> skip >> im >> skip, as qpp reads "(re,im)"
        char a = 0, b = 0, c = 0;
        double re = 0, im = 0;
        if (std::fscanf(f, " %c%lf %c%lf %c", &a, &re, &b, &im, &c) != 5) return false;
        z = cd(re, im);
        return true;
    };
    if (cols == 1) {
        std::vector<cd> h(std::min<uint64_t>(rows, kChunk));
        for (uint64_t e0 = 0; e0 < rows; e0 += h.size()) {
            const uint64_t n = std::min<uint64_t>(h.size(), rows - e0);
            for (uint64_t i = 0; i < n; ++i)
                if (!entry(h[i])) return qs::fail_runtime(QS_ERR_IO, E);
            if (int rc = store(st, e0, n, h.data(), tmp, W)) return rc;
        }
        return QS_OK;
    }
    std::vector<cd> h(rows * cols);
    for (uint64_t i = 0; i < rows; ++i)
        for (uint64_t j = 0; j < cols; ++j)
            if (!entry(h[i + j * rows])) return qs::fail_runtime(QS_ERR_IO, E);
    for (uint64_t e0 = 0; e0 < h.size(); e0 += kChunk) {
        const uint64_t n = std::min<uint64_t>(kChunk, h.size() - e0);
        if (int rc = store(st, e0, n, h.data() + e0, tmp, W)) return rc;
    }
    return QS_OK;
}

int load_binary(qs_state* st, std::FILE* f, uint64_t total, const char* W) {
    const char* E = "qpp::obsolete::load(): Input stream is corrupted!";
    std::vector<float> tmp;
    std::vector<cd> h(std::min<uint64_t>(total, kChunk));
    for (uint64_t e0 = 0; e0 < total; e0 += h.size()) {
        const uint64_t n = std::min<uint64_t>(h.size(), total - e0);
        if (std::fread(h.data(), sizeof(cd), n, f) != n) return qs::fail_runtime(QS_ERR_IO, E);
        if (int rc = store(st, e0, n, h.data(), tmp, W)) return rc;
    }
    return QS_OK;
}

}  // namespace

extern "C" {

int qs_state_save(const qs_state* st, const char* path, qs_format fmt) {
    constexpr const char* W = "qpp_amd::save()";
    if (!st || !st->d || !path) return fail(QS_ERR_INVALID_HANDLE, W);
    if (fmt != QS_FMT_TEXT && fmt != QS_FMT_BINARY) return fail(QS_ERR_OUT_OF_RANGE, W, "fmt");
    const uint64_t rows = st->D, cols = st->kind == QS_KET ? 1 : st->D;
    File f(std::fopen(path, "wb"));
    if (!f)
        return qs::fail_runtime(QS_ERR_IO, fmt == QS_FMT_TEXT
                                               ? "qpp::save(): Error writing output stream!"
                                               : "qpp::obsolete::save(): Error writing output stream!");
    QS_HIP(hipStreamSynchronize(st->stream), W);
    const int rc = fmt == QS_FMT_TEXT ? save_text(st, f.get(), rows, cols, W)
                                      : save_binary(st, f.get(), rows, cols, W);
    if (rc) return rc;
    if (std::fclose(f.release()) != 0)
        return qs::fail_runtime(QS_ERR_IO, "qpp::save(): Error writing output stream!");
    return QS_OK;
}

int qs_state_load(const char* path, qs_format fmt, qs_dtype dtype, qs_kind kind, uint32_t ndims,
                  const uint64_t* dims, qs_state** out) {
    constexpr const char* W = "qpp_amd::load()";
    if (!path || !out) return fail(QS_ERR_INVALID_HANDLE, W);
    if (fmt != QS_FMT_TEXT && fmt != QS_FMT_BINARY) return fail(QS_ERR_OUT_OF_RANGE, W, "fmt");
    File f(std::fopen(path, "rb"));
    if (!f)
        return qs::fail_runtime(QS_ERR_IO, fmt == QS_FMT_TEXT
                                               ? "qpp::load(): Error opening input stream!"
                                               : "qpp::obsolete::load(): Error opening input stream!");
    uint64_t rows = 0, cols = 0;
    if (fmt == QS_FMT_BINARY) {
        char hdr[kHeaderLen];
        size_t r = 0, c = 0;
        if (std::fread(hdr, 1, kHeaderLen, f.get()) != kHeaderLen ||
            std::memcmp(hdr, kHeader, kHeaderLen) != 0 || std::fread(&r, 8, 1, f.get()) != 1 ||
            std::fread(&c, 8, 1, f.get()) != 1)
            return qs::fail_runtime(QS_ERR_IO, "qpp::obsolete::load(): Input stream is corrupted!");
        rows = r;
        cols = c;
    } else {
        unsigned long long r = 0, c = 0;
        if (std::fscanf(f.get(), "%llu %llu", &r, &c) != 2)
            return qs::fail_runtime(QS_ERR_IO, "qpp::load(): Input stream is corrupted!");
        rows = r;
        cols = c;
    }
    qs_state* st = nullptr;
    if (int rc = qs_state_create(dtype, kind, ndims, dims, &st)) return rc;
    const bool shape_ok = rows == st->D && cols == (kind == QS_KET ? 1 : st->D);
    int rc = QS_OK;
    if (!shape_ok)
        rc = fail(kind == QS_KET ? QS_ERR_DIMS_MISMATCH_CVECTOR : QS_ERR_DIMS_MISMATCH_MATRIX, W,
                  "dims/file");
    else if (fmt == QS_FMT_TEXT)
        rc = load_text(st, f.get(), rows, cols, W);
    else
        rc = load_binary(st, f.get(), rows * cols, W);
    if (rc) {
        qs_state_destroy(st);
        return rc;
    }
    *out = st;
    return QS_OK;
}

}  // extern "C"
