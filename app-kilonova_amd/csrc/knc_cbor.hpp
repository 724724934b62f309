// knc_cbor.hpp -- CBOR (RFC 8949) writer and reader for the reference's
// checkpoint and products files.
//
// The reference writes them with serde through ciborium 0.1.0
// (src/io.rs:24-29, Cargo.lock:29-31), whose published design is to write
// every number in its shortest lossless form: integers in the smallest
// head, floats in the narrowest of binary16 / binary32 / binary64 that holds
// the value exactly (RFC 8949 section 4.2.2, "preferred serialization").
// Containers have definite lengths; structs are maps keyed by field name.
// The writer below produces exactly that; the reader accepts any well-formed
// CBOR of the expected shape (any float width, definite or indefinite
// containers, tags skipped), as serde's derived deserialisers do.
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace knc {
namespace cbor {

enum Major : uint8_t { UINT = 0, NINT = 1, BYTES = 2, TEXT = 3, ARRAY = 4, MAP = 5, TAG = 6, SIMPLE = 7 };

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline uint64_t bits_of(double x)
{
    uint64_t b;
    std::memcpy(&b, &x, 8);
    return b;
}

// Encodes x in its shortest exact float form into out; returns the byte
// count (3, 5 or 9).
inline int encode_f64(double x, uint8_t* out)
{
    const uint64_t b = bits_of(x);
    const uint64_t sign = b >> 63;
    const int bexp = (int)((b >> 52) & 0x7ff);
    const uint64_t man = b & ((1ull << 52) - 1);
    const uint64_t low42 = (1ull << 42) - 1;
    int half = -1;
    if (bexp == 0x7ff) {
        if ((man & low42) == 0) half = (int)((sign << 15) | 0x7c00u | (man >> 42));
    } else if (bexp == 0) {
        if (man == 0) half = (int)(sign << 15);   // +-0; binary64 subnormals are below binary16's range
    } else {
        const int e = bexp - 1023;
        if (e >= -14 && e <= 15) {
            if ((man & low42) == 0) half = (int)((sign << 15) | ((uint64_t)(e + 15) << 10) | (man >> 42));
        } else if (e >= -24 && e < -14) {
            // binary16 subnormal m 2^-24: the significand's bits below 2^-24 must be zero
            const int shift = 28 - e;
            const uint64_t sig = man | (1ull << 52);
            if ((sig & ((1ull << shift) - 1)) == 0) half = (int)((sign << 15) | (sig >> shift));
        }
    }
    if (half >= 0) {
        out[0] = 0xf9;
        out[1] = (uint8_t)(half >> 8);
        out[2] = (uint8_t)half;
        return 3;
    }
    const float f = (float)x;
    if (bits_of((double)f) == b) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        out[0] = 0xfa;
        for (int k = 0; k < 4; ++k) out[1 + k] = (uint8_t)(u >> (24 - 8 * k));
        return 5;
    }
    out[0] = 0xfb;
    for (int k = 0; k < 8; ++k) out[1 + k] = (uint8_t)(b >> (56 - 8 * k));
    return 9;
}

inline double decode_f16(uint16_t h)
{
    const int e = (h >> 10) & 0x1f, m = h & 0x3ff;
    double v;
    if (e == 0) v = std::ldexp((double)m, -24);
    else if (e == 31) {
        // NaN payloads keep their top bits, as a binary16 -> binary64 widening does
        const uint64_t b = 0x7ff0000000000000ull | ((uint64_t)m << 42);
        std::memcpy(&v, &b, 8);
    } else v = std::ldexp((double)(m | 0x400), e - 25);
    return (h & 0x8000) ? -v : v;
}

// Buffered writer into a FILE*.
class Writer {
public:
    explicit Writer(FILE* f) : f_(f), buf_(1u << 22) {}
    ~Writer() { flush(); }

    bool ok() const { return ok_; }

    void head(Major m, uint64_t n)
    {
        need(9);
        uint8_t* p = buf_.data() + pos_;
        const uint8_t mt = (uint8_t)(m << 5);
        if (n < 24) {
            p[0] = (uint8_t)(mt | n);
            pos_ += 1;
        } else if (n <= 0xff) {
            p[0] = (uint8_t)(mt | 24);
            p[1] = (uint8_t)n;
            pos_ += 2;
        } else if (n <= 0xffff) {
            p[0] = (uint8_t)(mt | 25);
            p[1] = (uint8_t)(n >> 8);
            p[2] = (uint8_t)n;
            pos_ += 3;
        } else if (n <= 0xffffffffull) {
            p[0] = (uint8_t)(mt | 26);
            for (int k = 0; k < 4; ++k) p[1 + k] = (uint8_t)(n >> (24 - 8 * k));
            pos_ += 5;
        } else {
            p[0] = (uint8_t)(mt | 27);
            for (int k = 0; k < 8; ++k) p[1 + k] = (uint8_t)(n >> (56 - 8 * k));
            pos_ += 9;
        }
    }
    void uint(uint64_t n) { head(UINT, n); }
    void sint(int64_t v)
    {
        if (v < 0) head(NINT, (uint64_t)(-1 - v));
        else head(UINT, (uint64_t)v);
    }
    void array(uint64_t n) { head(ARRAY, n); }
    void map(uint64_t n) { head(MAP, n); }
    void text(const char* s) { text(s, std::strlen(s)); }
    void text(const char* s, size_t n)
    {
        head(TEXT, n);
        raw((const uint8_t*)s, n);
    }
    void f64(double x)
    {
        need(9);
        pos_ += (size_t)encode_f64(x, buf_.data() + pos_);
    }
    // n floats, each preceded by nothing (elements of an enclosing array)
    void f64s(const double* x, size_t n)
    {
        for (size_t k = 0; k < n; ++k) f64(x[k]);
    }
    void raw(const uint8_t* p, size_t n)
    {
        if (n > buf_.size() - pos_) {
            flush();
            if (n > buf_.size()) {
                if (ok_ && std::fwrite(p, 1, n, f_) != n) ok_ = false;
                return;
            }
        }
        std::memcpy(buf_.data() + pos_, p, n);
        pos_ += n;
    }
    void flush()
    {
        if (pos_ && ok_ && std::fwrite(buf_.data(), 1, pos_, f_) != pos_) ok_ = false;
        pos_ = 0;
    }

private:
    void need(size_t n)
    {
        if (pos_ + n > buf_.size()) flush();
    }
    FILE* f_;
    std::vector<uint8_t> buf_;
    size_t pos_ = 0;
    bool ok_ = true;
};

// Cursor over an in-memory CBOR document.  Every read checks bounds and
// throws cbor::Error on malformed or unexpected input.
class Reader {
public:
    struct Head {
        uint8_t major;
        uint8_t minor;
        uint64_t arg;       // value, length, tag number or simple value
        bool indefinite;    // minor 31 on a string or container
    };

    Reader(const uint8_t* p, size_t n) : begin_(p), p_(p), end_(p + n) {}

    size_t offset() const { return (size_t)(p_ - begin_); }
    void seek(size_t off)
    {
        if (off > (size_t)(end_ - begin_)) throw Error("seek past the end");
        p_ = begin_ + off;
    }
    bool at_end() const { return p_ >= end_; }
    bool at_break() const { return p_ < end_ && *p_ == 0xff; }
    void take_break()
    {
        if (!at_break()) throw Error("expected a break");
        ++p_;
    }

    Head head()
    {
        if (p_ >= end_) throw Error("unexpected end of data");
        const uint8_t b = *p_++;
        Head h{(uint8_t)(b >> 5), (uint8_t)(b & 31), 0, false};
        if (h.minor < 24) {
            h.arg = h.minor;
        } else if (h.minor <= 27) {
            const int n = 1 << (h.minor - 24);
            if (end_ - p_ < n) throw Error("truncated head");
            for (int k = 0; k < n; ++k) h.arg = (h.arg << 8) | *p_++;
        } else if (h.minor == 31 && (h.major >= BYTES && h.major <= MAP)) {
            h.indefinite = true;
        } else if (h.minor == 31 && h.major == SIMPLE) {
            throw Error("unexpected break");
        } else {
            throw Error("reserved additional information");
        }
        return h;
    }

    // The head of the next data item, after any tags.
    Head item_head()
    {
        Head h = head();
        while (h.major == TAG) h = head();
        return h;
    }

    // Opens a container of the given major type; n = its length, or
    // UINT64_MAX when indefinite (iterate while !at_break(), then take_break()).
    uint64_t open(Major m)
    {
        const Head h = item_head();
        if (h.major != m) throw Error(m == MAP ? "expected a map" : "expected an array");
        return h.indefinite ? UINT64_MAX : h.arg;
    }
    bool more(uint64_t n, uint64_t k)
    {
        if (n != UINT64_MAX) return k < n;
        if (at_break()) {
            ++p_;
            return false;
        }
        return true;
    }

    double f64()
    {
        const Head h = item_head();
        if (h.major == SIMPLE) {
            if (h.minor == 25) return decode_f16((uint16_t)h.arg);
            if (h.minor == 26) {
                const uint32_t u = (uint32_t)h.arg;
                float f;
                std::memcpy(&f, &u, 4);
                return (double)f;
            }
            if (h.minor == 27) {
                double d;
                std::memcpy(&d, &h.arg, 8);
                return d;
            }
        } else if (h.major == UINT) {
            return (double)h.arg;
        } else if (h.major == NINT) {
            return -1.0 - (double)h.arg;
        }
        throw Error("expected a number");
    }

    int64_t sint()
    {
        const Head h = item_head();
        if (h.major == UINT && h.arg <= (uint64_t)INT64_MAX) return (int64_t)h.arg;
        if (h.major == NINT && h.arg <= (uint64_t)INT64_MAX) return -1 - (int64_t)h.arg;
        throw Error("expected an integer");
    }

    std::string text()
    {
        const Head h = item_head();
        if (h.major != TEXT) throw Error("expected a text string");
        std::string s;
        if (!h.indefinite) {
            take(s, h.arg);
            return s;
        }
        while (!at_break()) {
            const Head c = head();
            if (c.major != TEXT || c.indefinite) throw Error("bad text chunk");
            take(s, c.arg);
        }
        ++p_;
        return s;
    }

    // Skips one complete data item (nesting bounded by depth).
    void skip(int depth = 0)
    {
        if (depth > 64) throw Error("nesting too deep");
        const Head h = item_head();
        switch (h.major) {
        case UINT:
        case NINT:
        case SIMPLE:
            return;
        case BYTES:
        case TEXT:
            if (!h.indefinite) {
                advance(h.arg);
                return;
            }
            while (!at_break()) {
                const Head c = head();
                if (c.major != h.major || c.indefinite) throw Error("bad string chunk");
                advance(c.arg);
            }
            ++p_;
            return;
        case ARRAY:
        case MAP: {
            const uint64_t per = h.major == MAP ? 2 : 1;
            if (h.indefinite) {
                while (!at_break())
                    for (uint64_t q = 0; q < per; ++q) skip(depth + 1);
                ++p_;
                return;
            }
            for (uint64_t k = 0; k < h.arg; ++k)
                for (uint64_t q = 0; q < per; ++q) skip(depth + 1);
            return;
        }
        default:
            throw Error("unexpected item");
        }
    }

    // Reads n floats of a definite or indefinite array of exactly n numbers;
    // the common binary64 case is taken without the general path.
    void f64_array(uint64_t n_expected, double* out, size_t stride)
    {
        const uint64_t n = open(ARRAY);
        uint64_t k = 0;
        for (; more(n, k); ++k) {
            if (k >= n_expected) throw Error("array longer than its dimensions");
            if (p_ < end_ && *p_ == 0xfb && end_ - p_ >= 9) {
                uint64_t b = 0;
                for (int q = 1; q <= 8; ++q) b = (b << 8) | p_[q];
                p_ += 9;
                std::memcpy(&out[k * stride], &b, 8);
            } else {
                out[k * stride] = f64();
            }
        }
        if (k != n_expected) throw Error("array shorter than its dimensions");
    }

private:
    void advance(uint64_t n)
    {
        if ((uint64_t)(end_ - p_) < n) throw Error("truncated string");
        p_ += n;
    }
    void take(std::string& s, uint64_t n)
    {
        if ((uint64_t)(end_ - p_) < n) throw Error("truncated string");
        s.append((const char*)p_, (size_t)n);
        p_ += n;
    }

    const uint8_t* begin_;
    const uint8_t* p_;
    const uint8_t* end_;
};

}  // namespace cbor
}  // namespace knc
