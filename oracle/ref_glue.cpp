// ref_glue.cpp -- C entry points onto the reference's own compiled C++ code
// (lib/ldpc.cc, lib/matrices/*.cc, lib/ldpc_encoder.cc), for tests only.
// Built by oracle/Makefile into oracle/_ref/ from the sources where they lie;
// the reference decoder (lib/ldpc_decoder.cc) needs VOLK, which this image
// lacks, so it is not compiled here.
#include <stdint.h>
#include <string.h>
#include <exception>
#include "ldpc.h"
#include "ldpc_encoder.h"

using namespace wimax_ldpc_lib;

namespace {
// ldpc keeps its table and sizes protected (include/ldpc.h:38-48); a derived
// class may read them.
struct table_probe : public ldpc {
    table_probe(coderate r, unsigned z) : ldpc(r, z, 1) {}
    const int16_t* table() const { return m_checknode_array; }
    unsigned n() const { return m_N; }
    unsigned m() const { return m_M; }
    unsigned rows() const { return m_row_size; }
    unsigned cols() const { return m_col_size; }
    unsigned d() const { return m_D; }
    unsigned x() const { return m_X; }
};
}

extern "C" int ref_table_info(int rate, int z, int* n, int* m, int* rows, int* cols, int* d, int* x)
{
    try {
        table_probe p((coderate)rate, (unsigned)z);
        *n = p.n(); *m = p.m(); *rows = p.rows(); *cols = p.cols(); *d = p.d(); *x = p.x();
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

extern "C" int ref_table_copy(int rate, int z, int16_t* out)
{
    try {
        table_probe p((coderate)rate, (unsigned)z);
        memcpy(out, p.table(), sizeof(int16_t) * p.rows() * p.cols());
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

// Encode nframes frames with one reference encoder object.
extern "C" int ref_encode(int rate, int z, int nthreads, const uint8_t* info, uint8_t* cw, int nframes)
{
    try {
        ldpc_encoder enc((coderate)rate, (unsigned)z, (unsigned)nthreads);
        table_probe p((coderate)rate, (unsigned)z);
        const unsigned k = p.n() - p.m();
        for (int f = 0; f < nframes; f++)
            enc.encode_data(const_cast<uint8_t*>(info) + (size_t)f * k, cw + (size_t)f * p.n());
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}
