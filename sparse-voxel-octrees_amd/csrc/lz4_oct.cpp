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
#include <cstring>

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

inline void put_length(std::vector<uint8_t> &out, uint64_t extra) {
    while (extra >= 255) { out.push_back(255); extra -= 255; }
    out.push_back(uint8_t(extra));
}

}  // namespace

Lz4StreamEncoder::Lz4StreamEncoder() : base_(0), dictSize_(0), consumed_(0) {
    std::memset(table_, 0, sizeof(table_));
}

void Lz4StreamEncoder::encode(const uint8_t *src, uint64_t pos, uint32_t len, std::vector<uint8_t> &out) {
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

    auto put = [&](uint64_t p) { table_[hash_at(src + p)] = uint32_t(p - base_); };
    auto lookup = [&](uint32_t h) { return base_ + table_[h]; };

    auto emit_literals_token = [&](uint64_t ip) -> size_t {
        uint64_t lit = ip - anchor;
        size_t token = out.size();
        if (lit >= 15) {
            out.push_back(0xF0);
            put_length(out, lit - 15);
        } else {
            out.push_back(uint8_t(lit << 4));
        }
        out.insert(out.end(), src + anchor, src + ip);
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

            size_t token = emit_literals_token(ip);
            for (;;) {
                uint16_t off = uint16_t(ip - match);
                out.push_back(uint8_t(off));
                out.push_back(uint8_t(off >> 8));
                uint64_t ml = common_prefix(src + ip + kMinMatch, src + match + kMinMatch,
                                            matchEndLimit - (ip + kMinMatch));
                ip += kMinMatch + ml;
                if (ml >= 15) {
                    out[token] = uint8_t(out[token] + 15);
                    put_length(out, ml - 15);
                } else {
                    out[token] = uint8_t(out[token] + ml);
                }
                anchor = ip;
                if (ip > matchStartLimit) goto last_literals;
                put(ip - 2);
                match = lookup(hash_at(src + ip));
                put(ip);
                if (match + kWindow >= ip && load32(src + match) == load32(src + ip)) {
                    token = out.size();
                    out.push_back(0);
                    continue;  // immediate next match, no literals
                }
                break;
            }
            ip++;
            fwdHash = hash_at(src + ip);
        }
    }
last_literals : {
    uint64_t run = end - anchor;
    if (run >= 15) {
        out.push_back(0xF0);
        put_length(out, run - 15);
    } else {
        out.push_back(uint8_t(run << 4));
    }
    out.insert(out.end(), src + anchor, src + end);
}
    dictSize_ += len;
    consumed_ = end;
}

bool lz4_decode_block(const uint8_t *src, uint64_t srcLen, uint8_t *out, uint64_t pos, uint64_t outLen) {
    uint64_t ip = 0, op = pos;
    const uint64_t oend = pos + outLen;
    while (ip < srcLen) {
        uint32_t token = src[ip++];
        uint64_t lit = token >> 4;
        if (lit == 15) {
            uint8_t b;
            do {
                if (ip >= srcLen) return false;
                b = src[ip++];
                lit += b;
            } while (b == 255);
        }
        if (lit > srcLen - ip || lit > oend - op) return false;
        std::memcpy(out + op, src + ip, lit);
        ip += lit;
        op += lit;
        if (op == oend) return ip == srcLen;
        if (srcLen - ip < 2) return false;
        uint64_t off = uint64_t(src[ip]) | (uint64_t(src[ip + 1]) << 8);
        ip += 2;
        if (off == 0 || off > op) return false;
        uint64_t ml = token & 15u;
        if (ml == 15) {
            uint8_t b;
            do {
                if (ip >= srcLen) return false;
                b = src[ip++];
                ml += b;
            } while (b == 255);
        }
        ml += kMinMatch;
        if (ml > oend - op) return false;
        uint8_t *d = out + op;
        const uint8_t *s = d - off;
        if (off >= ml) {
            std::memcpy(d, s, ml);
        } else {
            for (uint64_t k = 0; k < ml; ++k) d[k] = s[k];
        }
        op += ml;
    }
    return op == oend;
}

int oct_decode(const uint8_t *data, uint64_t size, svo_octree &h) {
    if (size < 20) return fail(SVO_ERR_FORMAT, ".oct: truncated header");
    uint64_t n;
    std::memcpy(h.center, data, 12);
    std::memcpy(&n, data + 12, 8);
    if (n == 0 || n > (uint64_t(1) << 38)) return fail(SVO_ERR_FORMAT, ".oct: implausible word count");
    // every slice needs at least its 8-byte size field
    uint64_t slices = (n * 4 + kSlice - 1) / kSlice;
    if (slices * 8 > size - 20) return fail(SVO_ERR_FORMAT, ".oct: truncated (slice headers missing)");
    try {
        h.words.assign(n, 0);
    } catch (...) {
        return fail(SVO_ERR_NOMEM, ".oct: cannot allocate word array");
    }
    uint8_t *out = reinterpret_cast<uint8_t *>(h.words.data());
    uint64_t bytes = n * 4, p = 20;
    for (uint64_t off = 0; off < bytes; off += kSlice) {
        uint64_t cs;
        if (size - p < 8) return fail(SVO_ERR_FORMAT, ".oct: truncated slice header");
        std::memcpy(&cs, data + p, 8);
        p += 8;
        if (cs > size - p) return fail(SVO_ERR_FORMAT, ".oct: truncated slice");
        uint64_t os = bytes - off < kSlice ? bytes - off : kSlice;
        if (!lz4_decode_block(data + p, cs, out, off, os)) return fail(SVO_ERR_FORMAT, ".oct: corrupt LZ4 slice");
        p += cs;
    }
    return SVO_OK;
}

void oct_encode(const svo_octree &h, std::vector<uint8_t> &out) {
    uint64_t n = h.words.size(), bytes = n * 4;
    out.clear();
    out.reserve(20 + bytes / 2 + 1024);
    out.resize(20);
    std::memcpy(out.data(), h.center, 12);
    std::memcpy(out.data() + 12, &n, 8);
    const uint8_t *src = reinterpret_cast<const uint8_t *>(h.words.data());
    Lz4StreamEncoder enc;
    for (uint64_t off = 0; off < bytes; off += kSlice) {
        uint32_t len = uint32_t(bytes - off < kSlice ? bytes - off : kSlice);
        size_t at = out.size();
        out.resize(at + 8);
        enc.encode(src, off, len, out);
        uint64_t cs = out.size() - at - 8;
        std::memcpy(out.data() + at, &cs, 8);
    }
}

}  // namespace svo
