// Observation-table reader and period-file writers with the reference's
// file formats.
//
//   lsamd_read_table    replaces importObjXYData (source/main.cu:1781-1843):
//                       the file is a stream of numbers separated by commas
//                       and/or whitespace, taken in rows of 3 (objectId, t, y)
//                       or 4 (+ dy) values, every value parsed as the build's
//                       DTYPE (so ids pass through DTYPE, as in the reference)
//   lsamd_write_periods replaces outputPeriodsToFile (main.cu:961-976) and
//                       outputPeriodsToFileTopThree (main.cu:979-995):
//                       "id, period, power" or "id, p1, p2, p3, pow1, pow2,
//                       pow3" per object, 6 significant digits
//
// Parsing follows the reference's stream semantics: the text is split at
// commas; inside one comma-separated field, numbers separated by whitespace
// are read one after another until one fails to parse, and the rest of that
// field is skipped.  A number is an optional sign, digits with an optional
// fraction, and an optional exponent (what an istream reads for float or
// double; no "nan", "inf" or hexadecimal forms).

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "ls_internal.h"

namespace lsamd {
namespace {

bool read_file(const char *path, std::string &out) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long size = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize(size > 0 ? (size_t)size : 0);
    const size_t got = size > 0 ? std::fread(&out[0], 1, (size_t)size, f) : 0;
    std::fclose(f);
    return got == out.size();
}

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }
inline bool is_digit(char c) { return c >= '0' && c <= '9'; }

// Length of the number at p (0 if none): [+-]? digits [. digits] [eE [+-] digits],
// with at least one digit in the mantissa.
size_t number_length(const char *p, const char *end) {
    const char *q = p;
    if (q < end && (*q == '+' || *q == '-')) ++q;
    size_t mant = 0;
    while (q < end && is_digit(*q)) ++q, ++mant;
    if (q < end && *q == '.') {
        ++q;
        while (q < end && is_digit(*q)) ++q, ++mant;
    }
    if (mant == 0) return 0;
    if (q < end && (*q == 'e' || *q == 'E')) {
        const char *e = q + 1;
        if (e < end && (*e == '+' || *e == '-')) ++e;
        const char *d = e;
        while (e < end && is_digit(*e)) ++e;
        if (e > d) q = e;
    }
    return (size_t)(q - p);
}

// All values of the stream, parsed as R.
template <typename R>
void parse_values(const std::string &text, std::vector<R> &vals) {
    const char *p = text.data(), *end = p + text.size();
    char buf[128];
    while (p < end) {
        const char *fe = static_cast<const char *>(std::memchr(p, ',', (size_t)(end - p)));
        if (!fe) fe = end;
        const char *q = p;
        while (q < fe) {
            while (q < fe && is_space(*q)) ++q;
            if (q >= fe) break;
            const size_t len = number_length(q, fe);
            if (len == 0 || len >= sizeof(buf)) break;  // the field's stream fails here
            std::memcpy(buf, q, len);
            buf[len] = '\0';
            if constexpr (sizeof(R) == 4)
                vals.push_back(std::strtof(buf, nullptr));
            else
                vals.push_back(std::strtod(buf, nullptr));
            q += len;
        }
        p = fe + 1;
    }
}

template <typename R>
int read_table(const char *path, int with_dy, uint64_t capacity, uint32_t *ids, R *t, R *y, R *dy,
               uint64_t *rows) {
    std::string text;
    if (!read_file(path, text))
        return set_error(LSAMD_ERR_INVALID, std::string("lsamd_read_table: cannot read ") + path);
    std::vector<R> v;
    v.reserve(text.size() / 8);
    parse_values<R>(text, v);
    const unsigned cols = with_dy ? 4u : 3u;
    const uint64_t n = v.size() / cols;
    *rows = n;
    if (!ids) return LSAMD_OK;  // size query
    if (capacity < n)
        return set_error(LSAMD_ERR_INVALID, "lsamd_read_table: capacity " + std::to_string(capacity) +
                                                " < " + std::to_string(n) + " rows");
    if (!t || !y || (with_dy && !dy)) return set_error(LSAMD_ERR_INVALID, "lsamd_read_table: null output");
    for (uint64_t i = 0; i < n; ++i) {
        ids[i] = (uint32_t)v[i * cols + 0];
        t[i] = v[i * cols + 1];
        y[i] = v[i * cols + 2];
        if (with_dy) dy[i] = v[i * cols + 3];
    }
    return LSAMD_OK;
}

template <typename R>
int write_periods(const char *path, const uint32_t *ids, uint32_t K, uint32_t top_k, const R *per,
                  const R *pow) {
    std::ofstream out(path, std::ios::out);
    if (!out) return set_error(LSAMD_ERR_INVALID, std::string("lsamd_write_periods: cannot open ") + path);
    out.precision(6);
    for (uint64_t k = 0; k < K; ++k) {
        out << ids[k];
        if (top_k == 1) {
            out << ", " << per[k] << ", " << pow[k] << '\n';
        } else {
            const uint64_t o = k * 3;
            out << ", " << per[o] << ", " << per[o + 1] << ", " << per[o + 2] << ", " << pow[o] << ", "
                << pow[o + 1] << ", " << pow[o + 2] << '\n';
        }
    }
    out.close();
    if (!out) return set_error(LSAMD_ERR_INVALID, std::string("lsamd_write_periods: write failed: ") + path);
    return LSAMD_OK;
}

}  // namespace
}  // namespace lsamd

using namespace lsamd;

extern "C" {

int lsamd_read_table(const char *path, int dtype, int with_dy, uint64_t capacity, uint32_t *objectId,
                     void *t, void *y, void *dy, uint64_t *rows) {
    clear_error();
    if (!path || !rows) return set_error(LSAMD_ERR_INVALID, "lsamd_read_table: null path or rows");
    if (dtype == LSAMD_F32)
        return read_table<float>(path, with_dy, capacity, objectId, static_cast<float *>(t),
                                 static_cast<float *>(y), static_cast<float *>(dy), rows);
    if (dtype == LSAMD_F64)
        return read_table<double>(path, with_dy, capacity, objectId, static_cast<double *>(t),
                                  static_cast<double *>(y), static_cast<double *>(dy), rows);
    return set_error(LSAMD_ERR_INVALID, "lsamd_read_table: bad dtype");
}

int lsamd_write_periods(const char *path, const uint32_t *objectIds, uint32_t num_objects,
                        uint32_t top_k, int dtype, const void *periods, const void *powers) {
    clear_error();
    if (!path || !objectIds || !periods || !powers)
        return set_error(LSAMD_ERR_INVALID, "lsamd_write_periods: null argument");
    if (top_k != 1 && top_k != 3) return set_error(LSAMD_ERR_INVALID, "lsamd_write_periods: top_k must be 1 or 3");
    if (dtype == LSAMD_F32)
        return write_periods<float>(path, objectIds, num_objects, top_k, static_cast<const float *>(periods),
                                    static_cast<const float *>(powers));
    if (dtype == LSAMD_F64)
        return write_periods<double>(path, objectIds, num_objects, top_k, static_cast<const double *>(periods),
                                     static_cast<const double *>(powers));
    return set_error(LSAMD_ERR_INVALID, "lsamd_write_periods: bad dtype");
}

}  // extern "C"
