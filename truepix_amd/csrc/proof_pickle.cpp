// proof_pickle.cpp - protocol-5 pickle writer for TruePix proof objects (see
// proof_pickle.hpp for the rules it follows).  Host-only.
#include "proof_pickle.hpp"

#include <stdlib.h>
#include <string.h>

namespace tp {

namespace {

constexpr uint8_t OP_PROTO = 0x80, OP_FRAME = 0x95, OP_STOP = '.', OP_MARK = '(';
constexpr uint8_t OP_BININT = 'J', OP_BININT1 = 'K', OP_BININT2 = 'M', OP_LONG1 = 0x8a;
constexpr uint8_t OP_NONE = 'N', OP_SHORT_BINUNICODE = 0x8c, OP_BINUNICODE = 'X';
constexpr uint8_t OP_EMPTY_LIST = ']', OP_APPEND = 'a', OP_APPENDS = 'e';
constexpr uint8_t OP_EMPTY_TUPLE = ')', OP_TUPLE = 't', OP_TUPLE1 = 0x85, OP_TUPLE2 = 0x86,
                  OP_TUPLE3 = 0x87;
constexpr uint8_t OP_MEMOIZE = 0x94, OP_BINGET = 'h', OP_LONG_BINGET = 'j';
constexpr uint8_t OP_STACK_GLOBAL = 0x93, OP_NEWOBJ = 0x81, OP_BUILD = 'b';
constexpr size_t kFrameTarget = 64 * 1024, kFrameMin = 4, kFrameHeader = 9, kBatch = 1000;
constexpr uint64_t kP = (1ULL << 61) - 1;

// Encoded size of one int (opcode included) and the encoder itself.
size_t int_size(uint64_t v)
{
    if (v <= 0xff) return 2;
    if (v <= 0xffff) return 3;
    if (v <= 0x7fffffffULL) return 5;
    const int bits = 64 - __builtin_clzll(v);
    return 2 + (size_t)(bits / 8 + 1);
}

uint8_t* enc_int(uint8_t* d, uint64_t v)
{
    if (v <= 0xff) {
        d[0] = OP_BININT1;
        d[1] = (uint8_t)v;
        return d + 2;
    }
    if (v <= 0xffff) {
        d[0] = OP_BININT2;
        d[1] = (uint8_t)v;
        d[2] = (uint8_t)(v >> 8);
        return d + 3;
    }
    if (v <= 0x7fffffffULL) {
        d[0] = OP_BININT;
        for (int i = 0; i < 4; ++i) d[1 + i] = (uint8_t)(v >> (8 * i));
        return d + 5;
    }
    // LONG1: minimal two's-complement little-endian, bit_length / 8 + 1 bytes
    const int bits = 64 - __builtin_clzll(v);
    const int nb = bits / 8 + 1;
    d[0] = OP_LONG1;
    d[1] = (uint8_t)nb;
    for (int i = 0; i < nb; ++i) d[2 + i] = i < 8 ? (uint8_t)(v >> (8 * i)) : 0;
    return d + 2 + nb;
}

}  // namespace

PickleWriter::PickleWriter(std::vector<uint8_t>& o) : out(o) { out.clear(); }

void PickleWriter::start()
{
    put1(OP_PROTO);
    put1(5);
    framing = true;
}

void PickleWriter::finish()
{
    put1(OP_STOP);
    commit();
}

void PickleWriter::open_tuple()
{
    boundary();
    put1(OP_MARK);
}

void PickleWriter::close_tuple()
{
    put1(OP_TUPLE);
    memoize();
}

void PickleWriter::save(const PNode& x)
{
    boundary();
    switch (x.kind) {
    case PNode::INT: save_int(x.a); return;
    case PNode::NONE: put1(OP_NONE); return;
    case PNode::STR: save_str(x.s); return;
    case PNode::FP2: save_fp2(x.a, x.b); return;
    case PNode::LIST: save_list(x); return;
    case PNode::TUPLE: save_tuple(x); return;
    case PNode::INT_ARRAY: save_int_array(x.arr, x.n); return;
    case PNode::INT_ARRAY8: save_int_array(x.arr8, x.n); return;
    }
}

uint8_t* PickleWriter::grab(size_t n)
{
    if (framing && frame_start < 0) {
        frame_start = (long)out.size();
        out.resize(out.size() + kFrameHeader, 0xfe);
    }
    const size_t at = out.size();
    out.resize(at + n);
    return out.data() + at;
}

void PickleWriter::put(const void* p, size_t n) { memcpy(grab(n), p, n); }

void PickleWriter::commit()
{
    if (frame_start < 0) return;
    const size_t len = out.size() - (size_t)frame_start - kFrameHeader;
    if (len >= kFrameMin) {
        uint8_t* h = out.data() + frame_start;
        h[0] = OP_FRAME;
        for (int i = 0; i < 8; ++i) h[1 + i] = (uint8_t)((uint64_t)len >> (8 * i));
    } else {
        out.erase(out.begin() + frame_start, out.begin() + frame_start + kFrameHeader);
    }
    frame_start = -1;
}

void PickleWriter::boundary()
{
    if (frame_start >= 0 && out.size() - (size_t)frame_start - kFrameHeader >= kFrameTarget)
        commit();
}

void PickleWriter::memoize()
{
    put1(OP_MEMOIZE);
    ++memo;
}

void PickleWriter::get(uint32_t idx)
{
    if (idx < 256) {
        uint8_t b[2] = {OP_BINGET, (uint8_t)idx};
        put(b, 2);
    } else {
        uint8_t b[5] = {OP_LONG_BINGET, (uint8_t)idx, (uint8_t)(idx >> 8), (uint8_t)(idx >> 16),
                        (uint8_t)(idx >> 24)};
        put(b, 5);
    }
}

void PickleWriter::save_int(uint64_t v)
{
    uint8_t tmp[12];
    put(tmp, (size_t)(enc_int(tmp, v) - tmp));
}

void PickleWriter::save_str(const std::string& s)
{
    if (s.size() <= 0xff) {
        put1(OP_SHORT_BINUNICODE);
        put1((uint8_t)s.size());
    } else {
        put1(OP_BINUNICODE);
        uint8_t l[4];
        for (int i = 0; i < 4; ++i) l[i] = (uint8_t)(s.size() >> (8 * i));
        put(l, 4);
    }
    put(s.data(), s.size());
    memoize();
}

void PickleWriter::save_fp2(uint64_t re, uint64_t im)
{
    boundary();                       // save(cls)
    if (fp2_class < 0) {
        boundary();
        save_str("libTruePix.extfield");
        boundary();
        save_str("Fp2");
        put1(OP_STACK_GLOBAL);
        fp2_class = memo;
        memoize();
    } else {
        get((uint32_t)fp2_class);
    }
    boundary();                       // save(()) -> EMPTY_TUPLE, not memoized
    put1(OP_EMPTY_TUPLE);
    put1(OP_NEWOBJ);
    memoize();
    boundary();                       // save(state tuple)
    boundary();
    save_int(re);
    boundary();
    save_int(im);
    boundary();
    save_int(kP);
    put1(OP_TUPLE3);
    memoize();
    put1(OP_BUILD);
}

void PickleWriter::save_list(const PNode& x)
{
    put1(OP_EMPTY_LIST);
    memoize();
    const size_t n = x.kids.size();
    if (n == 0) return;
    if (n == 1) {
        save(x.kids[0]);
        put1(OP_APPEND);
        return;
    }
    size_t i = 0;
    do {
        put1(OP_MARK);
        for (size_t b = 0; b < kBatch && i < n; ++b, ++i) save(x.kids[i]);
        put1(OP_APPENDS);
    } while (i < n);
}

// A list of ints from an array (u64, or u8 for values below 256), batched
// and framed exactly like save_list.
template <class T>
void PickleWriter::save_int_array(const T* v, size_t n)
{
    put1(OP_EMPTY_LIST);
    memoize();
    if (n == 0) return;
    if (n == 1) {
        boundary();
        save_int(v[0]);
        put1(OP_APPEND);
        return;
    }
    size_t i = 0;
    do {
        put1(OP_MARK);
        const size_t end = i + kBatch < n ? i + kBatch : n;
        uint64_t big = 0;
        if (sizeof(T) > 1)
            for (size_t k = i; k < end; ++k) big |= v[k];
        const size_t worst = big <= 0xff ? 2 * (end - i) : [&] {
            size_t w = 0;
            for (size_t k = i; k < end; ++k) w += int_size(v[k]);
            return w;
        }();
        // No frame boundary falls inside this batch if the open frame cannot
        // reach the target before its last item starts.
        const size_t open = frame_start >= 0 ? out.size() - (size_t)frame_start - kFrameHeader : 0;
        if (frame_start >= 0 && open + worst < kFrameTarget) {
            uint8_t* d = grab(worst);
            if (big <= 0xff) {
                for (size_t k = i; k < end; ++k, d += 2) {
                    d[0] = OP_BININT1;
                    d[1] = (uint8_t)v[k];
                }
            } else {
                for (size_t k = i; k < end; ++k) d = enc_int(d, v[k]);
            }
            i = end;
        } else {
            for (; i < end; ++i) {
                boundary();
                save_int(v[i]);
            }
        }
        put1(OP_APPENDS);
    } while (i < n);
}

void PickleWriter::save_tuple(const PNode& x)
{
    const size_t n = x.kids.size();
    if (n == 0) {
        put1(OP_EMPTY_TUPLE);
        return;
    }
    if (n <= 3) {
        for (const PNode& k : x.kids) save(k);
        put1(n == 1 ? OP_TUPLE1 : n == 2 ? OP_TUPLE2 : OP_TUPLE3);
        memoize();
        return;
    }
    put1(OP_MARK);
    for (const PNode& k : x.kids) save(k);
    put1(OP_TUPLE);
    memoize();
}

void pickle_dumps(const PNode& root, std::vector<uint8_t>& out)
{
    PickleWriter w(out);
    w.start();
    w.save(root);
    w.finish();
}

namespace {

// Token form of a tree: 'i' u64 | 's' u32 len bytes | 'n' | 'f' u64 u64 |
// 'l' u32 count items | 't' u32 count items | 'a' u32 count u64 x count |
// 'b' u32 count u8 x count (a list of ints below 256) (little-endian).
bool decode(const uint8_t*& p, const uint8_t* end, PNode& x, int depth,
            std::vector<std::vector<uint64_t>>& store, std::vector<std::vector<uint8_t>>& store8)
{
    auto need = [&](size_t k) { return (size_t)(end - p) >= k; };
    auto rd32 = [&]() { uint32_t v; memcpy(&v, p, 4); p += 4; return v; };
    auto rd64 = [&]() { uint64_t v; memcpy(&v, p, 8); p += 8; return v; };
    if (depth > 64 || !need(1)) return false;
    const uint8_t t = *p++;
    switch (t) {
    case 'i':
        if (!need(8)) return false;
        x = PNode::integer(rd64());
        return true;
    case 'n': x = PNode::none(); return true;
    case 'f': {
        if (!need(16)) return false;
        const uint64_t re = rd64();
        x = PNode::fp2v(re, rd64());
        return true;
    }
    case 's': {
        if (!need(4)) return false;
        const uint32_t n = rd32();
        if (!need(n)) return false;
        x = PNode::str(std::string((const char*)p, n));
        p += n;
        return true;
    }
    case 'a': {
        if (!need(4)) return false;
        const uint32_t n = rd32();
        if (!need((size_t)n * 8)) return false;
        store.emplace_back(n);
        if (n) memcpy(store.back().data(), p, (size_t)n * 8);
        x = PNode::int_array(store.back().data(), n);
        p += (size_t)n * 8;
        return true;
    }
    case 'b': {
        if (!need(4)) return false;
        const uint32_t n = rd32();
        if (!need(n)) return false;
        store8.emplace_back(p, p + n);
        x = PNode::int_array8(store8.back().data(), n);
        p += n;
        return true;
    }
    case 'l':
    case 't': {
        if (!need(4)) return false;
        const uint32_t n = rd32();
        x = t == 'l' ? PNode::list() : PNode::tuple();
        x.kids.resize(n);
        for (uint32_t k = 0; k < n; ++k)
            if (!decode(p, end, x.kids[k], depth + 1, store, store8)) return false;
        return true;
    }
    default: return false;
    }
}

}  // namespace

}  // namespace tp

extern "C" {

// A15: pickle.dumps(tree, 5) of a token-encoded proof tree (token form above).
// *out is malloc'd; free it with tp_free.  0 / -1 on a malformed tree.
int tp_pickle_tree(const uint8_t* tokens, size_t n, uint8_t** out, size_t* out_len)
{
    const uint8_t* p = tokens;
    tp::PNode root;
    std::vector<std::vector<uint64_t>> store;
    std::vector<std::vector<uint8_t>> store8;
    if (!out || !out_len || !tp::decode(p, tokens + n, root, 0, store, store8) || p != tokens + n)
        return -1;
    std::vector<uint8_t> bytes;
    tp::pickle_dumps(root, bytes);
    *out = static_cast<uint8_t*>(malloc(bytes.size()));
    if (!*out) return -1;
    memcpy(*out, bytes.data(), bytes.size());
    *out_len = bytes.size();
    return 0;
}

void tp_free(void* p) { free(p); }

}  // extern "C"
