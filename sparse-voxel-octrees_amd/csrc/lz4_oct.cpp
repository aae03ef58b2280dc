// .oct container and its LZ4 stream codec.
//
// Container (src/VoxelOctree.cpp:57-123): float32 center[3], uint64 word
// count, then for every 64 MiB slice of the word bytes a uint64 compressed
// size followed by one LZ4 block. Slices form one LZ4 stream: a slice may
// copy from the 64 KiB that precede it.
//
// The encoder reproduces, byte for byte, what the reference's vendored LZ4
// 1.7.1 emits from LZ4_compress_continue (src/third-party/lz4.c:1011-1059,
// 442-650) on a 64-bit host: greedy single-probe matching over a 4096-entry
// hash table of 32-bit positions, 5-byte multiplicative hash
// ((read64 * 889523592379) >> 28) & 4095, skip acceleration (step grows by one
// every 64 failed probes), 64 KiB window, 12-byte end margin for match
// starts and 5 trailing literals, backward match extension, and the stream
// state carried across slices (positions relative to a moving base, the
// table rebased once the stream passes 2 GiB). These parameters fix the
// output, so they follow the reference; the code is written from the LZ4
// block format and that list of parameters.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <thread>

#include "svo_internal.hpp"

namespace svo {

namespace {

constexpr uint64_t kSlice = 64ull << 20;        // CompressionBlockSize, VoxelOctree.cpp:55
constexpr uint32_t kWindow = 65535;             // largest encodable match distance
constexpr uint32_t kMinMatch = 4;
constexpr uint32_t kEndLiterals = 5;            // last bytes are always literals
constexpr uint32_t kMatchStartMargin = 12;      // no match may start in the last 12 bytes
constexpr uint32_t kMinInput = kMatchStartMargin + 1;
constexpr uint64_t kRebaseAt = 0x80000000ull;   // table positions are rebased past 2 GiB

inline uint64_t load64(const uint8_t *p) { uint64_t v; std::memcpy(&v, p, 8); return v; }
inline uint32_t load32(const uint8_t *p) { uint32_t v; std::memcpy(&v, p, 4); return v; }

inline uint32_t hash_at(const uint8_t *p) {
    return uint32_t((load64(p) * 889523592379ull) >> 28) & 4095u;
}

// Number of equal bytes of a[0..) and b[0..), at most limit.
inline uint64_t common_prefix(const uint8_t *a, const uint8_t *b, uint64_t limit) {
    uint64_t n = 0;
    while (n + 8 <= limit) {
        uint64_t x = load64(a + n) ^ load64(b + n);
        if (x) return n + (uint64_t)(__builtin_ctzll(x) >> 3);
        n += 8;
    }
    while (n < limit && a[n] == b[n]) n++;
    return n;
}

inline uint8_t *put_length(uint8_t *o, uint64_t extra) {
    while (extra >= 255) {
        *o++ = 255;
        extra -= 255;
    }
    *o++ = uint8_t(extra);
    return o;
}

}  // namespace

Lz4StreamEncoder::Lz4StreamEncoder() : base_(0), dictSize_(0), consumed_(0) {
    std::memset(table_, 0, sizeof(table_));
}

uint64_t Lz4StreamEncoder::bound(uint32_t len) { return uint64_t(len) + len / 255 + 16; }

uint64_t Lz4StreamEncoder::encode(const uint8_t *src, uint64_t pos, uint32_t len, uint8_t *dst) {
    // Stream bookkeeping for a slice that directly follows the previous one.
    if (consumed_ - base_ > kRebaseAt) {
        uint64_t nb = consumed_ - 65536;
        for (uint32_t &e : table_) {
            uint64_t abs = base_ + e;
            e = abs < nb ? 0u : uint32_t(abs - nb);
        }
        base_ = nb;
        if (dictSize_ > 65536) dictSize_ = 65536;
    }
    const uint64_t lowLimit = pos - dictSize_;  // earliest byte a match may extend back to
    const uint64_t end = pos + len;
    const uint64_t matchStartLimit = end - kMatchStartMargin;
    const uint64_t matchEndLimit = end - kEndLiterals;
    uint64_t anchor = pos;
    uint8_t *o = dst;

    auto put = [&](uint64_t p) { table_[hash_at(src + p)] = uint32_t(p - base_); };
    auto lookup = [&](uint32_t h) { return base_ + table_[h]; };

    // literal run [anchor, ip) behind a new token; returns the token
    auto emit_literals = [&](uint64_t ip) -> uint8_t * {
        const uint64_t lit = ip - anchor;
        uint8_t *token = o++;
        if (lit >= 15) {
            *token = 0xF0;
            o = put_length(o, lit - 15);
        } else {
            *token = uint8_t(lit << 4);
        }
        std::memcpy(o, src + anchor, lit);
        o += lit;
        return token;
    };

    if (len >= kMinInput) {
        uint64_t ip = pos;
        put(ip);
        ip++;
        uint32_t fwdHash = hash_at(src + ip);
        for (;;) {
            // search a match, probing ever more sparsely
            uint64_t match;
            uint64_t fwd = ip;
            uint32_t step = 1, attempts = 1u << 6;
            bool tail = false;
            do {
                uint32_t h = fwdHash;
                ip = fwd;
                fwd += step;
                step = attempts++ >> 6;
                if (fwd > matchStartLimit) { tail = true; break; }
                match = lookup(h);
                fwdHash = hash_at(src + fwd);
                table_[h] = uint32_t(ip - base_);
            } while (match + kWindow < ip || load32(src + match) != load32(src + ip));
            if (tail) break;

            // extend backwards
            while (ip > anchor && match > lowLimit && src[ip - 1] == src[match - 1]) { ip--; match--; }

            uint8_t *token = emit_literals(ip);
            for (;;) {
                const uint16_t off = uint16_t(ip - match);
                *o++ = uint8_t(off);
                *o++ = uint8_t(off >> 8);
                uint64_t ml = common_prefix(src + ip + kMinMatch, src + match + kMinMatch,
                                            matchEndLimit - (ip + kMinMatch));
                ip += kMinMatch + ml;
                if (ml >= 15) {
                    *token = uint8_t(*token + 15);
                    o = put_length(o, ml - 15);
                } else {
                    *token = uint8_t(*token + ml);
                }
                anchor = ip;
                if (ip > matchStartLimit) goto last_literals;
                put(ip - 2);
                match = lookup(hash_at(src + ip));
                put(ip);
                if (match + kWindow >= ip && load32(src + match) == load32(src + ip)) {
                    token = o++;
                    *token = 0;
                    continue;  // immediate next match, no literals
                }
                break;
            }
            ip++;
            fwdHash = hash_at(src + ip);
        }
    }
last_literals : {
    const uint64_t run = end - anchor;
    if (run >= 15) {
        *o++ = 0xF0;
        o = put_length(o, run - 15);
    } else {
        *o++ = uint8_t(run << 4);
    }
    std::memcpy(o, src + anchor, run);
    o += run;
}
    dictSize_ += len;
    consumed_ = end;
    return uint64_t(o - dst);
}

namespace {

// Bytes of a slice whose value may depend on the previous slice's output
// while that output is not final yet (one bit per byte of the slice).
struct Taint {
    std::vector<uint64_t> bits;
    int64_t last = -1;  // last tainted byte, relative to the slice start
    explicit Taint(uint64_t n) : bits((n + 63) / 64, 0) {}
    bool get(uint64_t i) const { return (bits[i >> 6] >> (i & 63)) & 1u; }
    void set(uint64_t i) {
        bits[i >> 6] |= uint64_t(1) << (i & 63);
        last = int64_t(i);
    }
};

// Decodes one LZ4 block into out[pos, pos+outLen); matches may reach back
// into out[0, pos) (the previous slices of the stream). Every bound is
// checked.
//  - Wide: copies run 16 bytes at a time where the slice has room, which may
//    write a few bytes past the current position that later literals or
//    matches overwrite before anything reads them.
//  - taint != null: the bytes before pos are not final (another thread is
//    still writing them). They are not read; what they would feed is written
//    as 0 and marked tainted, transitively.
//  - stop >= 0: return after the token that writes byte pos+stop (the caller
//    re-decodes just the tainted prefix once the dictionary is final).
bool decode_slice(const uint8_t *src, uint64_t srcLen, uint8_t *out, uint64_t pos, uint64_t outLen, bool wide,
                  Taint *taint, int64_t stop) {
    const uint8_t *ip = src;
    const uint8_t *const iend = src + srcLen;
    uint8_t *const ostart = out + pos;
    uint8_t *op = ostart;
    uint8_t *const oend = op + outLen;
    while (ip < iend) {
        if (stop >= 0 && op - ostart > stop) return true;
        const uint32_t token = *ip++;
        uint64_t lit = token >> 4;
        if (lit == 15) {
            uint8_t b;
            do {
                if (ip >= iend) return false;
                b = *ip++;
                lit += b;
            } while (b == 255);
        }
        if (lit > uint64_t(iend - ip) || lit > uint64_t(oend - op)) return false;
        if (wide && uint64_t(iend - ip) >= lit + 16 && uint64_t(oend - op) >= lit + 16) {
            for (uint64_t k = 0; k < lit; k += 16) std::memcpy(op + k, ip + k, 16);
        } else {
            std::memcpy(op, ip, lit);
        }
        ip += lit;
        op += lit;
        if (op == oend) return ip == iend;
        if (iend - ip < 2) return false;
        const uint64_t off = uint64_t(ip[0]) | (uint64_t(ip[1]) << 8);
        ip += 2;
        if (off == 0 || off > uint64_t(op - out)) return false;
        uint64_t ml = token & 15u;
        if (ml == 15) {
            uint8_t b;
            do {
                if (ip >= iend) return false;
                b = *ip++;
                ml += b;
            } while (b == 255);
        }
        ml += kMinMatch;
        if (ml > uint64_t(oend - op)) return false;
        const uint8_t *s = op - off;
        if (taint && (s < ostart || int64_t(s - ostart) <= taint->last)) {
            // the source may lie in, or depend on, the unfinished dictionary
            for (uint64_t k = 0; k < ml; ++k) {
                const uint64_t d = uint64_t(op - ostart) + k;
                if (s + k < ostart) {
                    op[k] = 0;
                    taint->set(d);
                } else {
                    op[k] = s[k];
                    if (taint->get(uint64_t(s + k - ostart))) taint->set(d);
                }
            }
            op += ml;
            continue;
        }
        const bool room = wide && uint64_t(oend - op) >= ml + 16;
        if (room && off >= 16) {
            // each chunk reads bytes at least one chunk behind the ones it writes
            for (uint64_t k = 0; k < ml; k += 16) std::memcpy(op + k, s + k, 16);
        } else if (room && off >= 8) {
            for (uint64_t k = 0; k < ml; k += 8) std::memcpy(op + k, s + k, 8);
        } else if (ml <= 32) {
            for (uint64_t k = 0; k < ml; ++k) op[k] = s[k];
        } else if (off >= ml) {
            std::memcpy(op, s, ml);
        } else {
            // repeating pattern of period off: copy it with doubling strides,
            // every source range ending at or before its destination
            uint64_t done = 0, span = off;
            while (done < ml) {
                const uint64_t n = ml - done < span ? ml - done : span;
                std::memcpy(op + done, op + done - span, n);
                done += n;
                if (span <= done) span *= 2;
            }
        }
        op += ml;
    }
    return stop >= 0 ? true : op == oend;
}

struct SliceRef {
    uint64_t src, size, pos, len;
};

}  // namespace

bool lz4_decode_block(const uint8_t *src, uint64_t srcLen, uint8_t *out, uint64_t pos, uint64_t outLen) {
    return decode_slice(src, srcLen, out, pos, outLen, true, nullptr, -1);
}

// Slices decode in parallel: each slice first decodes without its dictionary
// (the 64 KiB before it), marking what depends on it; then, in stream order,
// only the marked prefix of each slice is decoded again against the final
// dictionary. Octree slices depend on their dictionary for a short prefix only.
int oct_decode(const uint8_t *data, uint64_t size, svo_octree &h) {
    if (size < 20) return fail(SVO_ERR_FORMAT, ".oct: truncated header");
    uint64_t n;
    std::memcpy(h.center, data, 12);
    std::memcpy(&n, data + 12, 8);
    if (n == 0 || n > (uint64_t(1) << 38)) return fail(SVO_ERR_FORMAT, ".oct: implausible word count");
    // every slice needs at least its 8-byte size field
    uint64_t slices = (n * 4 + kSlice - 1) / kSlice;
    if (slices * 8 > size - 20) return fail(SVO_ERR_FORMAT, ".oct: truncated (slice headers missing)");
    const uint64_t bytes = n * 4;
    std::vector<SliceRef> refs;
    uint64_t p = 20;
    for (uint64_t off = 0; off < bytes; off += kSlice) {
        uint64_t cs;
        if (size - p < 8) return fail(SVO_ERR_FORMAT, ".oct: truncated slice header");
        std::memcpy(&cs, data + p, 8);
        p += 8;
        if (cs > size - p) return fail(SVO_ERR_FORMAT, ".oct: truncated slice");
        const uint64_t len = bytes - off < kSlice ? bytes - off : kSlice;
        // An LZ4 sequence of k input bytes writes at most 255*k output bytes
        // (a match of 4 + 15 + 255*(e-1) + 254 bytes costs 3 + e bytes), so a
        // slice shorter than len/255 cannot decode to len bytes. Checking it
        // here bounds the word array by the file's size before allocating.
        if (len > 255 * cs) return fail(SVO_ERR_FORMAT, ".oct: slice too short for its decoded size");
        refs.push_back({p, cs, off, len});
        p += cs;
    }
    try {
        h.words.assign(n, 0);
    } catch (...) {
        return fail(SVO_ERR_NOMEM, ".oct: cannot allocate word array");
    }
    uint8_t *out = reinterpret_cast<uint8_t *>(h.words.data());
    const size_t ns = refs.size();
    std::vector<std::unique_ptr<Taint>> taint(ns);
    std::vector<char> ok(ns, 0);
    std::atomic<size_t> next{0};
    auto worker = [&]() {
        for (size_t k; (k = next.fetch_add(1)) < ns;) {
            const SliceRef &r = refs[k];
            if (k) taint[k].reset(new Taint(r.len));
            ok[k] = decode_slice(data + r.src, r.size, out, r.pos, r.len, true, taint[k].get(), -1);
        }
    };
    const size_t nt = std::min<size_t>(ns, std::max(1u, std::thread::hardware_concurrency()));
    if (nt <= 1) {
        worker();
    } else {
        std::vector<std::thread> pool;
        try {
            for (size_t t = 0; t < nt; ++t) pool.emplace_back(worker);
        } catch (...) {
        }
        worker();
        for (std::thread &t : pool) t.join();
    }
    for (size_t k = 0; k < ns; ++k)
        if (!ok[k]) return fail(SVO_ERR_FORMAT, ".oct: corrupt LZ4 slice");
    for (size_t k = 1; k < ns; ++k) {
        if (taint[k]->last < 0) continue;
        const SliceRef &r = refs[k];
        if (!decode_slice(data + r.src, r.size, out, r.pos, r.len, false, nullptr, taint[k]->last))
            return fail(SVO_ERR_FORMAT, ".oct: corrupt LZ4 slice");
    }
    return SVO_OK;
}

void oct_encode(const svo_octree &h, std::vector<uint8_t> &out) {
    uint64_t n = h.words.size(), bytes = n * 4;
    out.clear();
    uint64_t cap = 20;
    for (uint64_t off = 0; off < bytes; off += kSlice)
        cap += 8 + Lz4StreamEncoder::bound(uint32_t(bytes - off < kSlice ? bytes - off : kSlice));
    out.reserve(cap);
    out.resize(20);
    std::memcpy(out.data(), h.center, 12);
    std::memcpy(out.data() + 12, &n, 8);
    const uint8_t *src = reinterpret_cast<const uint8_t *>(h.words.data());
    Lz4StreamEncoder enc;
    for (uint64_t off = 0; off < bytes; off += kSlice) {
        uint32_t len = uint32_t(bytes - off < kSlice ? bytes - off : kSlice);
        size_t at = out.size();
        out.resize(at + 8 + Lz4StreamEncoder::bound(len));
        uint64_t cs = enc.encode(src, off, len, out.data() + at + 8);
        out.resize(at + 8 + cs);
        std::memcpy(out.data() + at, &cs, 8);
    }
}

}  // namespace svo
