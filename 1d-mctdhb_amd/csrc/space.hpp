// space.hpp - the device-resident configurational space (tables + LUTs).
#pragma once

#include <vector>

#include "fock.hpp"

namespace mctdhb {

constexpr int kMaxOrb = 30;    // reference MAX_ORBITALS (mctdhb_types.h:29)
constexpr int kMaxPar = 10000; // reference MAX_PARTICLES (mctdhb_types.h:31)
constexpr int kSpaceTables = 1;  // also build the reference-layout jump tables

struct Space {
    int npar = 0, norb = 0;
    uint32_t dim = 0;
    int device = 0;
    int ldw = 0;
    std::vector<uint32_t> w_host;       // rank LUT, (M-1) x ldw
    std::vector<uint32_t> subsdim_host; // reference subsdim layout
    uint32_t* d_W = nullptr;
    uint32_t* d_subsdim = nullptr;
    uint16_t* d_ht = nullptr;
    uint32_t* d_map = nullptr;
    uint32_t* d_mapot = nullptr;
    uint32_t* d_maptt = nullptr;
    uint32_t* d_strideot = nullptr;
    uint32_t* d_stridett = nullptr;
    uint64_t mapot_len = 0, maptt_len = 0;
    bool tables = false;
    // Work list of the paired H.C kernel (2 <= M): slots[0, n_pairs) are the
    // configurations i that start a pair (i, i + 1) of one run (fixed
    // n_2 .. n_{M-1}, ranks consecutive as one particle moves from orbital 0
    // to orbital 1), slots[n_pairs, n_pairs + n_singles) the configurations
    // left alone at the end of odd-length runs; all in rank order.
    // occupations as bytes with row stride 8 (built when M <= 8, N <= 255):
    // one 8-byte load per configuration in the fused kernels
    uint8_t* d_occ8 = nullptr;
    uint32_t* d_slots = nullptr;
    uint32_t n_pairs = 0, n_singles = 0;
    // complex64 staging with one zero element on each side (dim + 2), for
    // callers of the paired kernel whose vectors are not padded
    mutable void* d_pad = nullptr;
    // tiling of the mixed H.C kernel picked by timing both on this space:
    // -1 not yet timed, 0 one configuration per lane, 1 configuration pairs
    mutable int hc_tiling = -1;
};

int space_dimension_exact(int npar, int norb, uint32_t* dim);
int space_create(int npar, int norb, int flags, Space** out);
void space_destroy(Space* S);
uint32_t rank_host(int npar, int norb, const uint16_t* conf);
void unrank_host(uint32_t i, int npar, int norb, uint16_t* conf);

}  // namespace mctdhb
