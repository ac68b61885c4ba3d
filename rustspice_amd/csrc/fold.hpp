// fold.hpp - one workgroup of 256 threads folds a long list of terms into one
// value in the list's order (an ordered sum that must round as a serial
// loop does): wavefronts 1..3 compute the next kFoldChunk terms into LDS while
// thread 0 adds or subtracts the current ones one at a time (reads eight at
// a time, the operations serial: the chain of dependent FP64 additions is the
// bound, about 8 ns a term).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rsgpu {

constexpr int kFoldChunk = 1024;
// an RHS row with more contributions than this is assembled by a workgroup
constexpr int kFoldRhs = 4096;

// x (op) term(q0) (op) term(q0 + 1) ... (op) term(q1 - 1), op = - when kSub;
// the result is valid in thread 0. Every thread of the workgroup must call it.
template <bool kSub, class Term>
__device__ double fold_long(double x, int64_t q0, int64_t q1, Term term) {
    __shared__ double buf[2][kFoldChunk];
    const int tid = threadIdx.x, wave = tid / 64;
    for (int i = tid; i < kFoldChunk && q0 + i < q1; i += 256) buf[0][i] = term(q0 + i);
    __syncthreads();
    int k = 0;
    for (int64_t b = q0; b < q1; b += kFoldChunk, k++) {
        const int c = (int)((q1 - b) < kFoldChunk ? (q1 - b) : kFoldChunk), cur = k & 1;
        if (wave > 0) {
            const int64_t nb = b + kFoldChunk;
            for (int i = tid - 64; i < kFoldChunk && nb + i < q1; i += 192) buf[cur ^ 1][i] = term(nb + i);
        } else if (tid == 0) {
            const double* B = buf[cur];
            int i = 0;
            for (; i + 8 <= c; i += 8) {
                double t[8];
#pragma unroll
                for (int j = 0; j < 8; j++) t[j] = B[i + j];
#pragma unroll
                for (int j = 0; j < 8; j++) x = kSub ? x - t[j] : x + t[j];
            }
            for (; i < c; i++) x = kSub ? x - B[i] : x + B[i];
        }
        __syncthreads();
    }
    return x;
}

}  // namespace rsgpu
