// Encoder (aoa-ntt/include/newckks/ckks/encoder.h): canonical-embedding FFT
// (FP64, on the device), scaled truncation to residues, decoding from residue
// 0, and the slot rotations as Galois automorphisms.
#ifndef NEWCKKS_ENCODER_H
#define NEWCKKS_ENCODER_H

#include <complex>

#include "../arithmetic/poly_t.h"

class CKKSContext;

// vals[j] <- sum_s vals[s] w^(s 5^j), w = exp(2 pi i / (4 size)) (decoding
// direction), on host memory; runs on the device.
void fftSpecial(std::complex<double>* vals, const long size);
// The inverse of fftSpecial including the division by size (encoding direction).
void fftSpecialInv(std::complex<double>* vals, const long size);

void ckks_encode(CKKSContext* ctx, uint64_t* d_output, std::complex<double>* h_input, int filled_slots,
                 int empty_slots, uint64_t scalingfactor);
void ckks_decode(CKKSContext* ctx, std::complex<double>* h_val, uint64_t* d_a, int slots, uint64_t scalingfactor);
void rotate_slots_left(Context* ctx, poly_t* b, poly_t* a, int rotSlots);
void rotate_slots_right(Context* ctx, poly_t* b, poly_t* a, int rotSlots);
// In place, NTT domain afterwards (encoder.cu:404-427).
void conjugate_slots(Context* ctx, poly_t* a);

#endif
