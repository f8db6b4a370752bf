// Wire format for ciphertexts, plaintexts and keys.
//
// The reference has no serialization (its objects live in device memory and
// are never written out); this is the data-interchange format the drop-in adds.
// A blob is a 64-byte little-endian header followed by the object's uint64
// words exactly as they sit in device memory (limb-major, N coefficients per
// limb, NTT form where the object is):
//
//   off  size  field
//     0     4  magic 'K' 'B' 'F' 'V'
//     4     2  version (1)
//     6     2  kind (KBFV_WIRE_*)
//     8     4  log2 N
//    12     4  count: objects in the blob (a batch)
//    16    16  shape: 4 x uint32, the object's dimensions above N (unused = 1)
//    32     8  aux: Galois element (Galois keys), gadget base bits (RGSW), else 0
//    40     8  parameter fingerprint (FNV-1a 64 over the context's parameters)
//    48     8  payload bytes
//    56     8  payload checksum
//
// The checksum of words w_0 .. w_(n-1) is s0 ^ rotl(s1, 32) with s0 = sum w_i
// and s1 = sum (i + 1) w_i, both mod 2^64 (order-sensitive, vectorisable).
// Loading checks the header against the context, the checksum, and that every
// residue is below the modulus of its limb.
#include <cstring>
#include <vector>

#include "kbfv_internal.h"

namespace {

constexpr uint32_t WIRE_MAGIC = 0x5646424Bu;  // "KBFV"
constexpr uint16_t WIRE_VERSION = 1;
constexpr size_t WIRE_HEADER = 64;

struct Shape {
    uint32_t dims[4];   // outer dimensions of one object (product = limb rows per object)
    int limb_mod0;      // modulus index of limb 0 along the innermost limb dimension
    int limb_dim;       // number of limbs along that dimension (residues cycle with it)
    bool plain;         // residues mod the plaintext modulus (no limb structure)
};

uint64_t rows_of(const Shape &s) { return (uint64_t)s.dims[0] * s.dims[1] * s.dims[2] * s.dims[3]; }

int shape_of(const kbfv_context *c, int kind, uint64_t aux, Shape *s) {
    *s = Shape{{1, 1, 1, 1}, 0, 1, false};
    switch (kind) {
        case KBFV_WIRE_CIPHERTEXT:
        case KBFV_WIRE_PUBLIC_KEY:
            s->dims[0] = 2, s->dims[1] = (uint32_t)c->qn, s->limb_mod0 = c->oQ, s->limb_dim = c->qn;
            return KBFV_OK;
        case KBFV_WIRE_PLAINTEXT:
            s->plain = true;
            return KBFV_OK;
        case KBFV_WIRE_SECRET_KEY:
            s->dims[0] = (uint32_t)(c->pn + c->qn), s->limb_mod0 = c->oP, s->limb_dim = c->pn + c->qn;
            return KBFV_OK;
        case KBFV_WIRE_RELIN_KEY:
        case KBFV_WIRE_GALOIS_KEY:
            s->dims[0] = 2, s->dims[1] = (uint32_t)c->dnum, s->dims[2] = (uint32_t)c->ri_blocks;
            s->dims[3] = (uint32_t)c->tn, s->limb_mod0 = c->oT, s->limb_dim = c->tn;
            return KBFV_OK;
        case KBFV_WIRE_RGSW: {
            const int rows = kbfv_rgsw_rows(c, (int)aux);
            KB_REQUIRE(rows > 0, "bad gadget base for an RGSW blob");
            s->dims[0] = (uint32_t)rows, s->dims[1] = 2, s->dims[2] = (uint32_t)c->qn;
            s->limb_mod0 = c->oQ, s->limb_dim = c->qn;
            return KBFV_OK;
        }
        default:
            kb_set_error("unknown wire kind %d", kind);
            return KBFV_EINVAL;
    }
}

void put32(uint8_t *p, uint32_t v) { std::memcpy(p, &v, 4); }
void put64(uint8_t *p, uint64_t v) { std::memcpy(p, &v, 8); }
uint32_t get32(const uint8_t *p) { uint32_t v; std::memcpy(&v, p, 4); return v; }
uint64_t get64(const uint8_t *p) { uint64_t v; std::memcpy(&v, p, 8); return v; }

uint64_t fingerprint(const kbfv_context *c) {
    uint64_t h = 0xcbf29ce484222325ull;
    auto mix = [&](uint64_t v) {
        for (int b = 0; b < 8; b++) h = (h ^ ((v >> (8 * b)) & 0xff)) * 0x100000001b3ull;
    };
    mix((uint64_t)c->log_n), mix(c->plain), mix((uint64_t)c->qn), mix((uint64_t)c->pn), mix((uint64_t)c->tn);
    mix((uint64_t)c->rn), mix((uint64_t)c->gamma), mix((uint64_t)c->dnum), mix((uint64_t)c->ri_blocks);
    for (uint64_t q : c->mod) mix(q);
    return h;
}

}  // namespace

extern "C" uint64_t kbfv_wire_checksum(const uint64_t *words, size_t n) {
    uint64_t s0 = 0, s1 = 0;
    for (size_t i = 0; i < n; i++) s0 += words[i], s1 += (uint64_t)(i + 1) * words[i];
    return s0 ^ ((s1 << 32) | (s1 >> 32));
}

extern "C" size_t kbfv_wire_bytes(const kbfv_context *c, int kind, int count, uint64_t aux) {
    Shape s;
    if (!c || count <= 0 || shape_of(c, kind, aux, &s) != KBFV_OK) return 0;
    return WIRE_HEADER + (size_t)count * rows_of(s) * c->n * sizeof(uint64_t);
}

namespace {
// every residue below the modulus of its limb (rows cycle through the limb dimension)
int check_range(const kbfv_context *c, const Shape &s, const uint64_t *w, size_t words) {
    const size_t n = (size_t)c->n;
    const size_t rows = words / n;
    for (size_t r = 0; r < rows; r++) {
        const uint64_t q = s.plain ? c->plain : c->mod[s.limb_mod0 + (int)(r % (size_t)s.limb_dim)];
        const uint64_t *row = w + r * n;
        uint64_t bad = 0;
        for (size_t i = 0; i < n; i++) bad |= (uint64_t)(row[i] >= q);
        KB_REQUIRE(!bad, s.plain ? "wire payload residue out of range in row %zu (a plaintext travels as its "
                                   "coefficients mod t, the encode() form, not the encode_ntt form)"
                                 : "wire payload residue out of range in row %zu",
                   r);
    }
    return KBFV_OK;
}
}  // namespace

extern "C" int kbfv_wire_save(const kbfv_context *c, int kind, const uint64_t *dev, int count, uint64_t aux,
                              void *buf, size_t cap, void *stream) {
    KB_REQUIRE(c && dev && buf && count > 0, "bad wire-save arguments");
    Shape s;
    const int rc = shape_of(c, kind, aux, &s);
    if (rc != KBFV_OK) return rc;
    const size_t words = (size_t)count * rows_of(s) * c->n;
    KB_REQUIRE(cap >= WIRE_HEADER + words * 8, "wire buffer too small");
    uint8_t *p = (uint8_t *)buf;
    hipStream_t st = (hipStream_t)stream;
    KB_HIP(hipMemcpyAsync(p + WIRE_HEADER, dev, words * 8, hipMemcpyDeviceToHost, st));
    KB_HIP(hipStreamSynchronize(st));
    std::vector<uint64_t> tmp;
    const uint64_t *w = (const uint64_t *)(p + WIRE_HEADER);
    if (((uintptr_t)w & 7) != 0) {  // unaligned caller buffer: checksum a copy
        tmp.resize(words);
        std::memcpy(tmp.data(), w, words * 8);
        w = tmp.data();
    }
    // what load would refuse is refused here, before a blob exists
    KB_TRY(check_range(c, s, w, words));
    put32(p, WIRE_MAGIC);
    const uint16_t ver = WIRE_VERSION, k16 = (uint16_t)kind;
    std::memcpy(p + 4, &ver, 2), std::memcpy(p + 6, &k16, 2);
    put32(p + 8, (uint32_t)c->log_n), put32(p + 12, (uint32_t)count);
    for (int d = 0; d < 4; d++) put32(p + 16 + 4 * d, s.dims[d]);
    put64(p + 32, aux), put64(p + 40, fingerprint(c)), put64(p + 48, words * 8);
    put64(p + 56, kbfv_wire_checksum(w, words));
    return KBFV_OK;
}

extern "C" int kbfv_wire_peek(const void *buf, size_t len, int *kind, int *count, uint64_t *aux) {
    KB_REQUIRE(buf && len >= WIRE_HEADER, "wire blob shorter than its header");
    const uint8_t *p = (const uint8_t *)buf;
    KB_REQUIRE(get32(p) == WIRE_MAGIC, "not a kbfv wire blob (bad magic)");
    uint16_t ver, k16;
    std::memcpy(&ver, p + 4, 2), std::memcpy(&k16, p + 6, 2);
    KB_REQUIRE(ver == WIRE_VERSION, "unsupported wire version %u", (unsigned)ver);
    KB_REQUIRE(len >= WIRE_HEADER + get64(p + 48), "wire blob truncated");
    if (kind) *kind = k16;
    if (count) *count = (int)get32(p + 12);
    if (aux) *aux = get64(p + 32);
    return KBFV_OK;
}

extern "C" int kbfv_wire_load(const kbfv_context *c, const void *buf, size_t len, int kind, uint64_t *dev,
                              int count, void *stream) {
    KB_REQUIRE(c && buf && dev && count > 0, "bad wire-load arguments");
    int k = 0, cnt = 0;
    uint64_t aux = 0;
    int rc = kbfv_wire_peek(buf, len, &k, &cnt, &aux);
    if (rc != KBFV_OK) return rc;
    const uint8_t *p = (const uint8_t *)buf;
    KB_REQUIRE(k == kind, "wire blob holds kind %d, expected %d", k, kind);
    KB_REQUIRE(cnt == count, "wire blob holds %d objects, expected %d", cnt, count);
    KB_REQUIRE((int)get32(p + 8) == c->log_n, "wire blob is for log N = %u", get32(p + 8));
    KB_REQUIRE(get64(p + 40) == fingerprint(c), "wire blob was written under other parameters");
    Shape s;
    rc = shape_of(c, kind, aux, &s);
    if (rc != KBFV_OK) return rc;
    for (int d = 0; d < 4; d++) KB_REQUIRE(get32(p + 16 + 4 * d) == s.dims[d], "wire blob shape mismatch");
    const size_t words = (size_t)count * rows_of(s) * c->n;
    KB_REQUIRE(get64(p + 48) == words * 8, "wire payload size mismatch");
    std::vector<uint64_t> tmp;
    const uint64_t *w = (const uint64_t *)(p + WIRE_HEADER);
    if (((uintptr_t)w & 7) != 0) {
        tmp.resize(words);
        std::memcpy(tmp.data(), w, words * 8);
        w = tmp.data();
    }
    KB_REQUIRE(kbfv_wire_checksum(w, words) == get64(p + 56), "wire payload checksum mismatch");
    KB_TRY(check_range(c, s, w, words));
    hipStream_t st = (hipStream_t)stream;
    KB_HIP(hipMemcpyAsync(dev, w, words * 8, hipMemcpyHostToDevice, st));
    KB_HIP(hipStreamSynchronize(st));  // the caller's buffer may go away on return
    return KBFV_OK;
}
