// matrices.cpp -- the tables of include/matrices.h (the reference's
// lib/matrices/*.cc, one generated array per code), defined here for every
// (rate, z) and filled when the library is loaded, from the same expansion
// wl_code_table returns.  The decoder kernels do not read them: they are the
// reference's public data, for callers that inspect the codes.
#include <stddef.h>
#include <stdio.h>

#include "matrices.h"
#include "matrices.hpp"
#include "wimax_ldpc_amd.h"

#define WL_MATRIX(NAME, RATE, Z, ROWS, COLS) \
    unsigned int NAME##_D;                   \
    unsigned int NAME##_X;                   \
    int16_t NAME[ROWS][COLS];
#include "matrices_list.inc"
#undef WL_MATRIX

namespace {

void fill(int rate, int z, int16_t* table, size_t len, unsigned int* d, unsigned int* x)
{
    wl_code_info ci;
    if (wl_code_params(rate, z, &ci) != WL_OK || (size_t)ci.m * ci.c != len ||
        wl_code_table(rate, z, table, len) != WL_OK) {
        fprintf(stderr, "wimax_ldpc_amd: table of rate %d z %d: %s\n", rate, z, wl_last_error());
        return;
    }
    *d = (unsigned int)ci.d;
    *x = (unsigned int)ci.x;
}

struct Tables {
    Tables()
    {
#define WL_MATRIX(NAME, RATE, Z, ROWS, COLS) \
    fill(RATE, Z, &NAME[0][0], (size_t)(ROWS) * (COLS), &NAME##_D, &NAME##_X);
#include "matrices_list.inc"
#undef WL_MATRIX
    }
};

const Tables g_tables;

}  // namespace

int16_t* wl::matrix_table(int rate, int z)
{
#define WL_MATRIX(NAME, RATE, Z, ROWS, COLS) \
    if (rate == (RATE) && z == (Z)) return &NAME[0][0];
#include "matrices_list.inc"
#undef WL_MATRIX
    return nullptr;
}
