// proof_pickle.hpp - writes TruePix proof bytes from C++: the bytes that
// CPython's pickle.dumps(obj, protocol=5) produces for the proof object
// (TP/libTruePix/fiatshamir.py:234-237, with Fp2 objects as
// libTruePix.extfield.Fp2, TP/libTruePix/extfield.py:76-82).
//
// The proof holds only ints, strs, None, lists, tuples and Fp2 objects, none
// of them shared, so the writer follows CPython's C pickler for exactly
// these: every list, str, non-empty tuple and Fp2 is memoized in save order;
// lists append in batches of 1000 (a one-element list uses APPEND); ints use
// BININT1 / BININT2 / BININT up to 2^31 - 1 and LONG1 above; an Fp2 is
// (class) () NEWOBJ, then BUILD with the state tuple (real, imag, p), the
// class being STACK_GLOBAL the first time and a memo get afterwards.
// Framing (protocol >= 4): a frame opens at the first write after the
// previous one closed; at the start of saving any object the open frame is
// closed if it holds at least 64 KiB; a closed frame gets a FRAME header
// when it holds at least 4 bytes.  The final frame closes after STOP.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace tp {

struct PNode {
    enum Kind : uint8_t { INT, STR, NONE, FP2, LIST, TUPLE, INT_ARRAY, INT_ARRAY8 };
    Kind kind = NONE;
    uint64_t a = 0, b = 0;          // INT value; FP2 (real, imag)
    std::string s;                  // STR (utf-8)
    std::vector<PNode> kids;        // LIST / TUPLE items
    const uint64_t* arr = nullptr;  // INT_ARRAY: a list of ints held elsewhere
    const uint8_t* arr8 = nullptr;  // INT_ARRAY8: the same for ints below 256
    size_t n = 0;

    static PNode integer(uint64_t v) { PNode p; p.kind = INT; p.a = v; return p; }
    static PNode str(const std::string& v) { PNode p; p.kind = STR; p.s = v; return p; }
    static PNode none() { return PNode(); }
    static PNode fp2v(uint64_t re, uint64_t im) { PNode p; p.kind = FP2; p.a = re; p.b = im; return p; }
    static PNode list() { PNode p; p.kind = LIST; return p; }
    static PNode tuple() { PNode p; p.kind = TUPLE; return p; }
    static PNode int_array(const uint64_t* v, size_t n)
    {
        PNode p;
        p.kind = INT_ARRAY;
        p.arr = v;
        p.n = n;
        return p;
    }
    static PNode int_array8(const uint8_t* v, size_t n)
    {
        PNode p;
        p.kind = INT_ARRAY8;
        p.arr8 = v;
        p.n = n;
        return p;
    }
};

// pickle.dumps(root, 5) for the tree; the Fp2 modulus in every state tuple is
// p = 2^61 - 1.
void pickle_dumps(const PNode& root, std::vector<uint8_t>& out);

// The same writer, driven step by step, so a prefix of the proof (the
// top-level tuple's leading items) can be written before the rest is known:
//   start(); open_tuple(); save(a); save(b); ... close_tuple(); finish();
// produces pickle.dumps((a, b, ...), 5) for a tuple of more than 3 items.
class PickleWriter {
public:
    explicit PickleWriter(std::vector<uint8_t>& out);
    void start();
    void open_tuple();
    void close_tuple();
    void save(const PNode& x);
    void finish();

private:
    std::vector<uint8_t>& out;
    bool framing = false;
    long frame_start = -1;
    uint32_t memo = 0;
    long fp2_class = -1;

    uint8_t* grab(size_t n);
    void put1(uint8_t c) { *grab(1) = c; }
    void put(const void* p, size_t n);
    void commit();
    void boundary();
    void memoize();
    void get(uint32_t idx);
    void save_int(uint64_t v);
    void save_str(const std::string& s);
    void save_fp2(uint64_t re, uint64_t im);
    void save_list(const PNode& x);
    template <class T> void save_int_array(const T* v, size_t n);
    void save_tuple(const PNode& x);
};

}  // namespace tp
