// Internal declarations of the U-Net preconditioner kernels (dn_unet.hip)
// used by its host side (dn_unet.cpp).  Tensors are channels-last
// [n][H][W][C] BF16 (uint16_t bit patterns) unless stated otherwise.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace dn {

struct ConvArgs {
    int n, H, W, Ci, Co;       // Ci, Co multiples of 16
    const uint16_t *x;         // [n][H][W][Ci]
    const uint16_t *w;         // [3][3][Co][Ci]
    const float *bias;         // [Co] or null
    int relu;
    const uint16_t *mask;      // [n][H][W][Co] or null: output zeroed where mask <= 0
    uint16_t *y16;             // [n][H][W][Co] or null
    float *y32;                // [n][H][W][Co] fp32 or null
};
hipError_t launch_conv3x3(const ConvArgs &a, hipStream_t st);

struct WgradArgs {
    int n, H, W, Ci, Co;       // Ci <= 112, Co <= 64, multiples of 16
    const uint16_t *x;         // layer input [n][H][W][Ci]
    const uint16_t *dy;        // output gradient [n][H][W][Co]
    int chunk_rows;            // image rows per partial: wgrad_chunk_rows(n, H)
    float *part;               // [nchunks][9][Co][Ci]
    float *bpart;              // [nchunks][Co]
};
int wgrad_chunk_rows(int n, int H);
int wgrad_chunks(int n, int H);   // partials of a layer of n images of H rows
// dw [9][Co][Ci] and (if db) db [Co], fp32, summed in a fixed order
hipError_t launch_wgrad(const WgradArgs &a, float *dw, float *db, hipStream_t st);

hipError_t launch_pool(int n, int H, int W, int C, const uint16_t *x, uint16_t *y, hipStream_t st);
hipError_t launch_pool_bwd(int n, int H, int W, int C, const uint16_t *dpool, const uint16_t *dcat, int ccat, int off,
                           const uint16_t *act, uint16_t *dact, hipStream_t st);
hipError_t launch_upcat(int n, int H, int W, int Cu, int Cs, const uint16_t *up, const uint16_t *skip, uint16_t *cat,
                        hipStream_t st);
hipError_t launch_up_bwd(int n, int H, int W, int Cu, int Ccat, const uint16_t *dcat, const uint16_t *act_low,
                         uint16_t *dlow, hipStream_t st);
hipError_t launch_pack_input(int n, int H, int W, const float *x, uint16_t *xp, hipStream_t st);
hipError_t launch_mse(long long P, const float *y, const float *t, uint16_t *dy, double *part, int nblocks,
                      hipStream_t st);
hipError_t launch_extract(long long P, const float *y, float *out, hipStream_t st);
hipError_t launch_cast_weights(int Co, int Ci, const float *w32, uint16_t *w16, uint16_t *wt16, hipStream_t st);
hipError_t launch_adam(long long n, float *w, const float *g, float *m, float *v, float lr_t, float b1, float b2,
                       float eps, hipStream_t st);

}  // namespace dn
