// matrices_check.cpp -- a caller written against the reference's public
// headers: it names the generated tables of include/matrices.h directly (as
// code using the reference's installed headers does), reads them through the
// ldpc class, and writes every table with its D and X for the Python test to
// compare with the reference's own tables (tests/golden/tables_ref.json).
//   matrices_check OUT.bin
// OUT.bin: per code, int32 {rate, z, rows, cols, D, X}, then rows*cols int16.
#include <stdio.h>
#include <stdint.h>
#include <string.h>

#include "BER_runner.h"
#include "ldpc.h"

namespace {

// ldpc exposes m_checknode_array to derived classes (include/ldpc.h), as the
// reference's decoder and encoder use it
struct table_view : wimax_ldpc_lib::ldpc {
    table_view(coderate r, unsigned z) : ldpc(r, z, 1) {}
    const int16_t* table() const { return m_checknode_array; }
    unsigned rows() const { return m_row_size; }
    unsigned cols() const { return m_col_size; }
    unsigned d() const { return m_D; }
    unsigned x() const { return m_X; }
};

int fails = 0;

void expect(bool ok, const char* what)
{
    if (!ok) {
        printf("FAIL %s\n", what);
        fails++;
    }
}

void dump(FILE* f, int rate, int z, int rows, int cols, unsigned d, unsigned x, const int16_t* t)
{
    const int32_t hdr[6] = {rate, z, rows, cols, (int32_t)d, (int32_t)x};
    fwrite(hdr, sizeof hdr, 1, f);
    fwrite(t, sizeof(int16_t), (size_t)rows * cols, f);
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 2) {
        printf("usage: %s OUT.bin\n", argv[0]);
        return 2;
    }
#if !defined(Z_24) || !defined(Z_96)
    expect(false, "Z_24 / Z_96 gates");
#endif
    // the tables by name
    expect(wimax_2304_0_83_D == 80 && wimax_2304_0_83_X == 1, "wimax_2304_0_83 D/X");
    expect(wimax_2304_0_83[0][0] == 1 && wimax_2304_0_83[0][19] == 2016, "wimax_2304_0_83 row 0");
    expect(wimax_576_0_5[0][5] >= 0 && wimax_576_0_5[0][6] == -1, "wimax_576_0_5 row 0 (degree 6, -1 tail)");
    expect(sizeof(wimax_576_0_5) == 288 * 7 * sizeof(int16_t), "wimax_576_0_5 size");
    expect(sizeof(wimax_1536_0_75B) / sizeof(wimax_1536_0_75B[0]) == 384, "wimax_1536_0_75B rows");
    // the code object points at the same table
    {
        table_view v(FIVESIXTHS, 96);
        expect(v.table() == &wimax_2304_0_83[0][0], "ldpc::m_checknode_array is wimax_2304_0_83");
        expect(v.rows() == 384 && v.cols() == 20 && v.d() == wimax_2304_0_83_D && v.x() == wimax_2304_0_83_X,
               "ldpc members");
    }
    // a runner without a code: counters at zero, frames refused
    {
        wimax_ldpc_lib::BER_runner none;
        expect(none.get_total_num_bits() == 0 && none.get_total_num_errors() == 0, "BER_runner() counters");
        bool threw = false;
        try {
            none.run_iteration(3.0);
        } catch (const std::exception&) {
            threw = true;
        }
        expect(threw, "BER_runner() run_iteration throws");
    }
    FILE* f = fopen(argv[1], "wb");
    if (!f) {
        printf("cannot write %s\n", argv[1]);
        return 2;
    }
    static const coderate rates[6] = {HALFRATE, TWOTHIRDSA, TWOTHIRDSB, THREEQUARTERSA, THREEQUARTERSB, FIVESIXTHS};
    for (int r = 0; r < 6; r++)
        for (unsigned z = 24; z <= 96; z += 4) {
            table_view v(rates[r], z);
            dump(f, r, (int)z, (int)v.rows(), (int)v.cols(), v.d(), v.x(), v.table());
        }
    fclose(f);
    printf("%s\n", fails ? "FAILED" : "ok");
    return fails ? 1 : 0;
}
