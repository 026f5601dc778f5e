// Shared host/device encodings of the engine: lowered device ops, item
// descriptors and the tableau layout.
//
// Tableau layout of one item (one circuit, S shots sharing the X/Z blocks):
//   generators are COLUMNS (sdim/tableau/dataclasses.py:14-15): stabilizers
//   0..n-1 then destabilizers n..2n-1, G = 2n columns; rows are qudits.
//   odd d : X[n][rs], Z[n][rs] uint8 (one generator per byte, value in [0,d)),
//           Ph[S][rs] uint8 phases mod d;  rs = round_up(G, 16) bytes.
//   d = 2 : X[n][rs], Z[n][rs] bit-packed (generator g = bit g%64 of u64 word
//           g/64), Ph[S][2][rs] two bit-planes p0,p1 with phase = p0 + 2 p1
//           (mod 4);  rs = 8 * round_up(ceil(G/64), 2) bytes.
//   The X and Z blocks are qudit-major, rows padded to np = round_up(n, 8), with
//   an optional 8-row interleave (il = 3; il = 0 is plain row-major): byte b of
//   row r sits at cell(r, b, rs, il) = ((r >> il) rs + 16 (b/16)) 2^il +
//   16 (r mod 2^il) + b%16.  With il = 3 a 128-byte line holds 16 bytes of 8
//   consecutive qudits, so reading one generator column (the pivot of a
//   measurement) touches n/8 lines instead of n; items whose gates run straight
//   from HBM (gang members, the plain path) keep il = 0.  Z starts np rows after
//   X (Z rows are X rows np.. of one block); the phase vectors follow, row-major.
//   Bytes of an item: T = 2 np rs + S rs (odd) or 2 np rs + 2 S rs (d = 2).
// The X/Z blocks are identical for every shot of a circuit (only the phases
// depend on the random outcomes and noise draws), so an item stores them once
// and keeps one phase vector per shot.
#pragma once
#include <stdint.h>

namespace sdim {

// Tableaux of at most this many qudits run serial gate layers with their rows in registers
// (one-wave workgroups of the launch class of small tableaux).
constexpr uint32_t SDIM_TINY_ROWS = 16;
// ... when an item holds at most this many shots (the class's one wave also runs the per-shot
// phase passes)
constexpr uint32_t SDIM_TINY_SHOTS = 64;

enum OpKind : uint32_t {
  OP_H = 0, OP_HI = 1, OP_P = 2, OP_PI = 3, OP_SUM = 4, OP_SUMI = 5,
  OP_PAULI = 6,     // X^a Z^b conjugation, a = byte1, b = byte2 (shot-invariant)
  OP_NOISE = 7,     // N1: channel = byte1, keep53 index = t, noise slot in segment = byte2|byte3<<8
  OP_MEAS = 8,      // t = measurement slot
  OP_RESET = 9,     // t = measurement slot; then X^((-m) mod d)
  OP_SEG = 10,      // gate segment header: q = #layers, t = #noise ops in the segment
  OP_LAYER = 11,    // layer header: q = #ops in the layer; t = 1: a serial layer (ops in program
                    // order that may share qudits, applied one after another)
  OP_END = 12,
  // closed forms of the reference's words (sdim/tableau/tableau_gates.py), equal on every column:
  OP_CZ = 13,       // H^-1(t) SUM(c,t) H(t):    z[t] += x[c], z[c] += x[t], phase += po x[c] x[t]
  OP_CZI = 14,      // H^-1(t) SUM^-1(c,t) H(t): z[t] -= x[c], z[c] -= x[t], phase -= po x[c] x[t]
  OP_SWAP = 15,     // the nine-primitive prime word: rows q and t exchanged, phases unchanged
  OP_SNAP = 16,     // copy the item's tableau to snapshot slot t (after instruction t of the circuit)
};

struct DevOp {
  uint32_t w0;   // kind | a << 8 | b << 16 | c << 24
  uint32_t q;
  uint32_t t;
  uint32_t src;  // instruction index in the user's circuit (Philox counter word 0)
};

static inline __host__ __device__ uint32_t op_kind(const DevOp& o) { return o.w0 & 0xFFu; }
static inline __host__ __device__ uint32_t op_a(const DevOp& o) { return (o.w0 >> 8) & 0xFFu; }
static inline __host__ __device__ uint32_t op_b(const DevOp& o) { return (o.w0 >> 16) & 0xFFu; }
static inline __host__ __device__ uint32_t op_c(const DevOp& o) { return (o.w0 >> 24) & 0xFFu; }
static inline __host__ __device__ uint32_t op_noise_slot(const DevOp& o) { return (o.w0 >> 16) & 0xFFFFu; }

// Ops of the Pauli-frame sampler (frame.hip): a DevOp whose kind is the gate id
// (sdim/gatedata.py:42-102) of a gate that changes the frame; X, Z and their
// inverses are dropped (they leave a frame unchanged).  Gates: q, t.  M, M_X,
// RESET: byte a = the reference shot's record (value | det << 7), t = event
// index.  N1: bits 8-9 = channel, bits 10-31 = keep53 >> 32, t = keep53 low
// word.  src = instruction index (Philox counter word 0).
enum FrameOpKind : uint32_t {
  FOP_H = 5, FOP_HI = 6, FOP_P = 7, FOP_PI = 8, FOP_CNOT = 9, FOP_CNOTI = 10, FOP_CZ = 11,
  FOP_CZI = 12, FOP_SWAP = 13, FOP_M = 14, FOP_MX = 15, FOP_RESET = 16, FOP_NOISE = 17,
};

enum ItemFlags : uint32_t {
  ITEM_LDS = 1u,        // tableau resident in LDS for the whole item
  ITEM_INIT_GIVEN = 2u, // start from the tableau already stored at tab_off (else identity)
  ITEM_PH_LDS = 4u,     // X/Z in HBM, the shots' phase vectors in LDS for the whole item
  ITEM_SNAP_MEAS = 8u,  // copy the tableau to snapshot slot e right after measurement event e
                        // (before RESET's fix-up); the snapshots start at snap_off
  // frame items: ITEM_LDS = the frames of the item's shots live in LDS (else at tab_off)
};

struct ItemDesc {
  uint64_t tab_off;    // byte offset of the item's tableau in the arena (final state lands here)
  uint64_t rec_off;    // byte offset of records: [n_meas][S] uint8 = value | det << 7
  uint64_t seed;       // Philox key word 0
  uint64_t stream;     // Philox key word 1
  uint64_t keep_off;   // index of the circuit's first keep53 threshold
  uint32_t ops_off;    // first DevOp of the circuit in the op pool
  uint32_t n_ops;
  uint32_t n;
  uint32_t d;
  uint32_t S;          // shots in this item
  uint32_t shot0;      // global index of the first shot (Philox counter word 1)
  uint32_t rs;         // row stride in bytes
  uint32_t flags;
  uint32_t n_meas;
  uint32_t il;         // X/Z row interleave: blocks of 2^il rows (0: row-major)
  uint32_t id;         // index of the item in its submission (reported with kernel errors)
  uint64_t snap_off;   // byte offset of the item's tableau snapshots (ITEM_SNAP_MEAS, OP_SNAP), each
                       // snapshot_stride() bytes in the item's own layout
};

// bytes between two snapshots of an item's tableau
static inline __host__ __device__ uint64_t snapshot_stride(uint32_t n, uint32_t d, uint32_t S);

static inline __host__ __device__ uint32_t row_stride(uint32_t n, uint32_t d) {
  uint32_t G = 2 * n;
  if (d == 2) {
    uint32_t w = (G + 63) / 64;
    w = (w + 1) & ~1u;
    return 8 * w;
  }
  return (G + 15) & ~15u;
}

static inline __host__ __device__ uint32_t rows_padded(uint32_t n) { return (n + 7u) & ~7u; }

// byte offset of byte b of row r in an X/Z block of row stride rs with row interleave il
static inline __host__ __device__ uint64_t cell(uint32_t r, uint32_t b, uint32_t rs, uint32_t il) {
  return (((uint64_t)(r >> il) * rs + (b & ~15u)) << il) + (r & ((1u << il) - 1u)) * 16u + (b & 15u);
}

static inline __host__ __device__ uint64_t tableau_bytes(uint32_t n, uint32_t d, uint32_t S) {
  uint64_t rs = row_stride(n, d);
  return 2ull * rows_padded(n) * rs + (d == 2 ? 2ull : 1ull) * S * rs;
}

static inline __host__ __device__ uint64_t snapshot_stride(uint32_t n, uint32_t d, uint32_t S) {
  return (tableau_bytes(n, d, S) + 255u) & ~255ull;
}

}  // namespace sdim
