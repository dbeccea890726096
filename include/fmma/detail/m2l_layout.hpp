// m2l_layout.hpp -- element order of the M2L operator fragments.
//
// The M2L contraction reads each offset's P x P operator K[kl][km] as MFMA
// A operands, padded to MP rows (16 * ceil(P/16)) and KP columns.  Element r
// of one offset's block of MP * KP holds K[row][col] (zero outside P x P):
//   FMMA_FRAG_F64 / FMMA_FRAG_F32 (16x16x4):
//       r = (rb * (KP/4) + s) * 64 + lane,      row = 16 rb + lane % 16, col = 4 s + lane / 16
//   FMMA_FRAG_BF16 (16x16x32, hi and lo parts in two arrays):
//       r = ((rb * (KP/32) + s) * 64 + lane) * 8 + j,  row = 16 rb + lane % 16, col = 32 s + 8 (lane / 16) + j
//   FMMA_FRAG_F32W (32x32x2, four steps per lane side by side for one 16-byte load; MP = 128):
//       r = ((rb * (KP/8) + g) * 64 + lane) * 4 + j,   row = 32 rb + lane % 32, col = 2 (4 g + j) + lane / 32
//   FMMA_FRAG_BF16W (32x32x16, hi and lo parts in two arrays; MP = 128):
//       r = ((rb * (KP/16) + s) * 64 + lane) * 8 + j,  row = 32 rb + lane % 32, col = 16 s + 8 (lane / 32) + j
#pragma once
#include "../../fmma_hip.h"

namespace fmma {
namespace detail {

__host__ __device__ inline void frag_rowcol(int format, int KP, long r, int &row, int &col)
{
    if (format == FMMA_FRAG_BF16W) {
        const int j = (int)(r & 7), lane = (int)((r >> 3) & 63);
        const long q = r >> 9;
        const int s = (int)(q % (KP / 16)), rb = (int)(q / (KP / 16));
        row = 32 * rb + (lane & 31);
        col = 16 * s + 8 * (lane >> 5) + j;
    } else if (format == FMMA_FRAG_F32W) {
        const int j = (int)(r & 3), lane = (int)((r >> 2) & 63);
        const long q = r >> 8;
        const int g = (int)(q % (KP / 8)), rb = (int)(q / (KP / 8));
        row = 32 * rb + (lane & 31);
        col = 2 * (4 * g + j) + (lane >> 5);
    } else if (format == FMMA_FRAG_BF16) {
        const int j = (int)(r & 7), lane = (int)((r >> 3) & 63);
        const long q = r >> 9;
        const int s = (int)(q % (KP / 32)), rb = (int)(q / (KP / 32));
        row = 16 * rb + (lane & 15);
        col = 32 * s + 8 * (lane >> 4) + j;
    } else {
        const int lane = (int)(r & 63);
        const long q = r >> 6;
        const int s = (int)(q % (KP / 4)), rb = (int)(q / (KP / 4));
        row = 16 * rb + (lane & 15);
        col = 4 * s + (lane >> 4);
    }
}

}  // namespace detail
}  // namespace fmma
