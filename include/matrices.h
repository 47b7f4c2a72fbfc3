/* matrices.h -- the expanded 802.16e parity-check tables, drop-in for the
 * reference's include/matrices.h.
 *
 * wimax_<n>_<rate>[M][C]: row m lists the variables of check m in ascending
 * order, -1 after the last one (n = 24 z, M = Mb z).  _D is the shift of the
 * top nonzero block of the first parity column at this z and _X the base row
 * of its middle (shift 0) block, the encoder constants.  libwimax_ldpc_amd.so
 * defines all of them and fills them when it is loaded.
 * Written by wimax_ldpc_lib_amd/csrc/gen_matrices.py. */
#ifndef MATRICES_H
#define MATRICES_H

#include <stdint.h>

#define Z_24
#define Z_28
#define Z_32
#define Z_36
#define Z_40
#define Z_44
#define Z_48
#define Z_52
#define Z_56
#define Z_60
#define Z_64
#define Z_68
#define Z_72
#define Z_76
#define Z_80
#define Z_84
#define Z_88
#define Z_92
#define Z_96

#ifdef Z_24
extern unsigned int wimax_576_0_5_D;
extern unsigned int wimax_576_0_5_X;
extern int16_t wimax_576_0_5[288][7];
#endif

#ifdef Z_24
extern unsigned int wimax_576_0_66A_D;
extern unsigned int wimax_576_0_66A_X;
extern int16_t wimax_576_0_66A[192][10];
#endif

#ifdef Z_24
extern unsigned int wimax_576_0_66B_D;
extern unsigned int wimax_576_0_66B_X;
extern int16_t wimax_576_0_66B[192][11];
#endif

#ifdef Z_24
extern unsigned int wimax_576_0_75A_D;
extern unsigned int wimax_576_0_75A_X;
extern int16_t wimax_576_0_75A[144][15];
#endif

#ifdef Z_24
extern unsigned int wimax_576_0_75B_D;
extern unsigned int wimax_576_0_75B_X;
extern int16_t wimax_576_0_75B[144][15];
#endif

#ifdef Z_24
extern unsigned int wimax_576_0_83_D;
extern unsigned int wimax_576_0_83_X;
extern int16_t wimax_576_0_83[96][20];
#endif

#ifdef Z_28
extern unsigned int wimax_672_0_5_D;
extern unsigned int wimax_672_0_5_X;
extern int16_t wimax_672_0_5[336][7];
#endif

#ifdef Z_28
extern unsigned int wimax_672_0_66A_D;
extern unsigned int wimax_672_0_66A_X;
extern int16_t wimax_672_0_66A[224][10];
#endif

#ifdef Z_28
extern unsigned int wimax_672_0_66B_D;
extern unsigned int wimax_672_0_66B_X;
extern int16_t wimax_672_0_66B[224][11];
#endif

#ifdef Z_28
extern unsigned int wimax_672_0_75A_D;
extern unsigned int wimax_672_0_75A_X;
extern int16_t wimax_672_0_75A[168][15];
#endif

#ifdef Z_28
extern unsigned int wimax_672_0_75B_D;
extern unsigned int wimax_672_0_75B_X;
extern int16_t wimax_672_0_75B[168][15];
#endif

#ifdef Z_28
extern unsigned int wimax_672_0_83_D;
extern unsigned int wimax_672_0_83_X;
extern int16_t wimax_672_0_83[112][20];
#endif

#ifdef Z_32
extern unsigned int wimax_768_0_5_D;
extern unsigned int wimax_768_0_5_X;
extern int16_t wimax_768_0_5[384][7];
#endif

#ifdef Z_32
extern unsigned int wimax_768_0_66A_D;
extern unsigned int wimax_768_0_66A_X;
extern int16_t wimax_768_0_66A[256][10];
#endif

#ifdef Z_32
extern unsigned int wimax_768_0_66B_D;
extern unsigned int wimax_768_0_66B_X;
extern int16_t wimax_768_0_66B[256][11];
#endif

#ifdef Z_32
extern unsigned int wimax_768_0_75A_D;
extern unsigned int wimax_768_0_75A_X;
extern int16_t wimax_768_0_75A[192][15];
#endif

#ifdef Z_32
extern unsigned int wimax_768_0_75B_D;
extern unsigned int wimax_768_0_75B_X;
extern int16_t wimax_768_0_75B[192][15];
#endif

#ifdef Z_32
extern unsigned int wimax_768_0_83_D;
extern unsigned int wimax_768_0_83_X;
extern int16_t wimax_768_0_83[128][20];
#endif

#ifdef Z_36
extern unsigned int wimax_864_0_5_D;
extern unsigned int wimax_864_0_5_X;
extern int16_t wimax_864_0_5[432][7];
#endif

#ifdef Z_36
extern unsigned int wimax_864_0_66A_D;
extern unsigned int wimax_864_0_66A_X;
extern int16_t wimax_864_0_66A[288][10];
#endif

#ifdef Z_36
extern unsigned int wimax_864_0_66B_D;
extern unsigned int wimax_864_0_66B_X;
extern int16_t wimax_864_0_66B[288][11];
#endif

#ifdef Z_36
extern unsigned int wimax_864_0_75A_D;
extern unsigned int wimax_864_0_75A_X;
extern int16_t wimax_864_0_75A[216][15];
#endif

#ifdef Z_36
extern unsigned int wimax_864_0_75B_D;
extern unsigned int wimax_864_0_75B_X;
extern int16_t wimax_864_0_75B[216][15];
#endif

#ifdef Z_36
extern unsigned int wimax_864_0_83_D;
extern unsigned int wimax_864_0_83_X;
extern int16_t wimax_864_0_83[144][20];
#endif

#ifdef Z_40
extern unsigned int wimax_960_0_5_D;
extern unsigned int wimax_960_0_5_X;
extern int16_t wimax_960_0_5[480][7];
#endif

#ifdef Z_40
extern unsigned int wimax_960_0_66A_D;
extern unsigned int wimax_960_0_66A_X;
extern int16_t wimax_960_0_66A[320][10];
#endif

#ifdef Z_40
extern unsigned int wimax_960_0_66B_D;
extern unsigned int wimax_960_0_66B_X;
extern int16_t wimax_960_0_66B[320][11];
#endif

#ifdef Z_40
extern unsigned int wimax_960_0_75A_D;
extern unsigned int wimax_960_0_75A_X;
extern int16_t wimax_960_0_75A[240][15];
#endif

#ifdef Z_40
extern unsigned int wimax_960_0_75B_D;
extern unsigned int wimax_960_0_75B_X;
extern int16_t wimax_960_0_75B[240][15];
#endif

#ifdef Z_40
extern unsigned int wimax_960_0_83_D;
extern unsigned int wimax_960_0_83_X;
extern int16_t wimax_960_0_83[160][20];
#endif

#ifdef Z_44
extern unsigned int wimax_1056_0_5_D;
extern unsigned int wimax_1056_0_5_X;
extern int16_t wimax_1056_0_5[528][7];
#endif

#ifdef Z_44
extern unsigned int wimax_1056_0_66A_D;
extern unsigned int wimax_1056_0_66A_X;
extern int16_t wimax_1056_0_66A[352][10];
#endif

#ifdef Z_44
extern unsigned int wimax_1056_0_66B_D;
extern unsigned int wimax_1056_0_66B_X;
extern int16_t wimax_1056_0_66B[352][11];
#endif

#ifdef Z_44
extern unsigned int wimax_1056_0_75A_D;
extern unsigned int wimax_1056_0_75A_X;
extern int16_t wimax_1056_0_75A[264][15];
#endif

#ifdef Z_44
extern unsigned int wimax_1056_0_75B_D;
extern unsigned int wimax_1056_0_75B_X;
extern int16_t wimax_1056_0_75B[264][15];
#endif

#ifdef Z_44
extern unsigned int wimax_1056_0_83_D;
extern unsigned int wimax_1056_0_83_X;
extern int16_t wimax_1056_0_83[176][20];
#endif

#ifdef Z_48
extern unsigned int wimax_1152_0_5_D;
extern unsigned int wimax_1152_0_5_X;
extern int16_t wimax_1152_0_5[576][7];
#endif

#ifdef Z_48
extern unsigned int wimax_1152_0_66A_D;
extern unsigned int wimax_1152_0_66A_X;
extern int16_t wimax_1152_0_66A[384][10];
#endif

#ifdef Z_48
extern unsigned int wimax_1152_0_66B_D;
extern unsigned int wimax_1152_0_66B_X;
extern int16_t wimax_1152_0_66B[384][11];
#endif

#ifdef Z_48
extern unsigned int wimax_1152_0_75A_D;
extern unsigned int wimax_1152_0_75A_X;
extern int16_t wimax_1152_0_75A[288][15];
#endif

#ifdef Z_48
extern unsigned int wimax_1152_0_75B_D;
extern unsigned int wimax_1152_0_75B_X;
extern int16_t wimax_1152_0_75B[288][15];
#endif

#ifdef Z_48
extern unsigned int wimax_1152_0_83_D;
extern unsigned int wimax_1152_0_83_X;
extern int16_t wimax_1152_0_83[192][20];
#endif

#ifdef Z_52
extern unsigned int wimax_1248_0_5_D;
extern unsigned int wimax_1248_0_5_X;
extern int16_t wimax_1248_0_5[624][7];
#endif

#ifdef Z_52
extern unsigned int wimax_1248_0_66A_D;
extern unsigned int wimax_1248_0_66A_X;
extern int16_t wimax_1248_0_66A[416][10];
#endif

#ifdef Z_52
extern unsigned int wimax_1248_0_66B_D;
extern unsigned int wimax_1248_0_66B_X;
extern int16_t wimax_1248_0_66B[416][11];
#endif

#ifdef Z_52
extern unsigned int wimax_1248_0_75A_D;
extern unsigned int wimax_1248_0_75A_X;
extern int16_t wimax_1248_0_75A[312][15];
#endif

#ifdef Z_52
extern unsigned int wimax_1248_0_75B_D;
extern unsigned int wimax_1248_0_75B_X;
extern int16_t wimax_1248_0_75B[312][15];
#endif

#ifdef Z_52
extern unsigned int wimax_1248_0_83_D;
extern unsigned int wimax_1248_0_83_X;
extern int16_t wimax_1248_0_83[208][20];
#endif

#ifdef Z_56
extern unsigned int wimax_1344_0_5_D;
extern unsigned int wimax_1344_0_5_X;
extern int16_t wimax_1344_0_5[672][7];
#endif

#ifdef Z_56
extern unsigned int wimax_1344_0_66A_D;
extern unsigned int wimax_1344_0_66A_X;
extern int16_t wimax_1344_0_66A[448][10];
#endif

#ifdef Z_56
extern unsigned int wimax_1344_0_66B_D;
extern unsigned int wimax_1344_0_66B_X;
extern int16_t wimax_1344_0_66B[448][11];
#endif

#ifdef Z_56
extern unsigned int wimax_1344_0_75A_D;
extern unsigned int wimax_1344_0_75A_X;
extern int16_t wimax_1344_0_75A[336][15];
#endif

#ifdef Z_56
extern unsigned int wimax_1344_0_75B_D;
extern unsigned int wimax_1344_0_75B_X;
extern int16_t wimax_1344_0_75B[336][15];
#endif

#ifdef Z_56
extern unsigned int wimax_1344_0_83_D;
extern unsigned int wimax_1344_0_83_X;
extern int16_t wimax_1344_0_83[224][20];
#endif

#ifdef Z_60
extern unsigned int wimax_1440_0_5_D;
extern unsigned int wimax_1440_0_5_X;
extern int16_t wimax_1440_0_5[720][7];
#endif

#ifdef Z_60
extern unsigned int wimax_1440_0_66A_D;
extern unsigned int wimax_1440_0_66A_X;
extern int16_t wimax_1440_0_66A[480][10];
#endif

#ifdef Z_60
extern unsigned int wimax_1440_0_66B_D;
extern unsigned int wimax_1440_0_66B_X;
extern int16_t wimax_1440_0_66B[480][11];
#endif

#ifdef Z_60
extern unsigned int wimax_1440_0_75A_D;
extern unsigned int wimax_1440_0_75A_X;
extern int16_t wimax_1440_0_75A[360][15];
#endif

#ifdef Z_60
extern unsigned int wimax_1440_0_75B_D;
extern unsigned int wimax_1440_0_75B_X;
extern int16_t wimax_1440_0_75B[360][15];
#endif

#ifdef Z_60
extern unsigned int wimax_1440_0_83_D;
extern unsigned int wimax_1440_0_83_X;
extern int16_t wimax_1440_0_83[240][20];
#endif

#ifdef Z_64
extern unsigned int wimax_1536_0_5_D;
extern unsigned int wimax_1536_0_5_X;
extern int16_t wimax_1536_0_5[768][7];
#endif

#ifdef Z_64
extern unsigned int wimax_1536_0_66A_D;
extern unsigned int wimax_1536_0_66A_X;
extern int16_t wimax_1536_0_66A[512][10];
#endif

#ifdef Z_64
extern unsigned int wimax_1536_0_66B_D;
extern unsigned int wimax_1536_0_66B_X;
extern int16_t wimax_1536_0_66B[512][11];
#endif

#ifdef Z_64
extern unsigned int wimax_1536_0_75A_D;
extern unsigned int wimax_1536_0_75A_X;
extern int16_t wimax_1536_0_75A[384][15];
#endif

#ifdef Z_64
extern unsigned int wimax_1536_0_75B_D;
extern unsigned int wimax_1536_0_75B_X;
extern int16_t wimax_1536_0_75B[384][15];
#endif

#ifdef Z_64
extern unsigned int wimax_1536_0_83_D;
extern unsigned int wimax_1536_0_83_X;
extern int16_t wimax_1536_0_83[256][20];
#endif

#ifdef Z_68
extern unsigned int wimax_1632_0_5_D;
extern unsigned int wimax_1632_0_5_X;
extern int16_t wimax_1632_0_5[816][7];
#endif

#ifdef Z_68
extern unsigned int wimax_1632_0_66A_D;
extern unsigned int wimax_1632_0_66A_X;
extern int16_t wimax_1632_0_66A[544][10];
#endif

#ifdef Z_68
extern unsigned int wimax_1632_0_66B_D;
extern unsigned int wimax_1632_0_66B_X;
extern int16_t wimax_1632_0_66B[544][11];
#endif

#ifdef Z_68
extern unsigned int wimax_1632_0_75A_D;
extern unsigned int wimax_1632_0_75A_X;
extern int16_t wimax_1632_0_75A[408][15];
#endif

#ifdef Z_68
extern unsigned int wimax_1632_0_75B_D;
extern unsigned int wimax_1632_0_75B_X;
extern int16_t wimax_1632_0_75B[408][15];
#endif

#ifdef Z_68
extern unsigned int wimax_1632_0_83_D;
extern unsigned int wimax_1632_0_83_X;
extern int16_t wimax_1632_0_83[272][20];
#endif

#ifdef Z_72
extern unsigned int wimax_1728_0_5_D;
extern unsigned int wimax_1728_0_5_X;
extern int16_t wimax_1728_0_5[864][7];
#endif

#ifdef Z_72
extern unsigned int wimax_1728_0_66A_D;
extern unsigned int wimax_1728_0_66A_X;
extern int16_t wimax_1728_0_66A[576][10];
#endif

#ifdef Z_72
extern unsigned int wimax_1728_0_66B_D;
extern unsigned int wimax_1728_0_66B_X;
extern int16_t wimax_1728_0_66B[576][11];
#endif

#ifdef Z_72
extern unsigned int wimax_1728_0_75A_D;
extern unsigned int wimax_1728_0_75A_X;
extern int16_t wimax_1728_0_75A[432][15];
#endif

#ifdef Z_72
extern unsigned int wimax_1728_0_75B_D;
extern unsigned int wimax_1728_0_75B_X;
extern int16_t wimax_1728_0_75B[432][15];
#endif

#ifdef Z_72
extern unsigned int wimax_1728_0_83_D;
extern unsigned int wimax_1728_0_83_X;
extern int16_t wimax_1728_0_83[288][20];
#endif

#ifdef Z_76
extern unsigned int wimax_1824_0_5_D;
extern unsigned int wimax_1824_0_5_X;
extern int16_t wimax_1824_0_5[912][7];
#endif

#ifdef Z_76
extern unsigned int wimax_1824_0_66A_D;
extern unsigned int wimax_1824_0_66A_X;
extern int16_t wimax_1824_0_66A[608][10];
#endif

#ifdef Z_76
extern unsigned int wimax_1824_0_66B_D;
extern unsigned int wimax_1824_0_66B_X;
extern int16_t wimax_1824_0_66B[608][11];
#endif

#ifdef Z_76
extern unsigned int wimax_1824_0_75A_D;
extern unsigned int wimax_1824_0_75A_X;
extern int16_t wimax_1824_0_75A[456][15];
#endif

#ifdef Z_76
extern unsigned int wimax_1824_0_75B_D;
extern unsigned int wimax_1824_0_75B_X;
extern int16_t wimax_1824_0_75B[456][15];
#endif

#ifdef Z_76
extern unsigned int wimax_1824_0_83_D;
extern unsigned int wimax_1824_0_83_X;
extern int16_t wimax_1824_0_83[304][20];
#endif

#ifdef Z_80
extern unsigned int wimax_1920_0_5_D;
extern unsigned int wimax_1920_0_5_X;
extern int16_t wimax_1920_0_5[960][7];
#endif

#ifdef Z_80
extern unsigned int wimax_1920_0_66A_D;
extern unsigned int wimax_1920_0_66A_X;
extern int16_t wimax_1920_0_66A[640][10];
#endif

#ifdef Z_80
extern unsigned int wimax_1920_0_66B_D;
extern unsigned int wimax_1920_0_66B_X;
extern int16_t wimax_1920_0_66B[640][11];
#endif

#ifdef Z_80
extern unsigned int wimax_1920_0_75A_D;
extern unsigned int wimax_1920_0_75A_X;
extern int16_t wimax_1920_0_75A[480][15];
#endif

#ifdef Z_80
extern unsigned int wimax_1920_0_75B_D;
extern unsigned int wimax_1920_0_75B_X;
extern int16_t wimax_1920_0_75B[480][15];
#endif

#ifdef Z_80
extern unsigned int wimax_1920_0_83_D;
extern unsigned int wimax_1920_0_83_X;
extern int16_t wimax_1920_0_83[320][20];
#endif

#ifdef Z_84
extern unsigned int wimax_2016_0_5_D;
extern unsigned int wimax_2016_0_5_X;
extern int16_t wimax_2016_0_5[1008][7];
#endif

#ifdef Z_84
extern unsigned int wimax_2016_0_66A_D;
extern unsigned int wimax_2016_0_66A_X;
extern int16_t wimax_2016_0_66A[672][10];
#endif

#ifdef Z_84
extern unsigned int wimax_2016_0_66B_D;
extern unsigned int wimax_2016_0_66B_X;
extern int16_t wimax_2016_0_66B[672][11];
#endif

#ifdef Z_84
extern unsigned int wimax_2016_0_75A_D;
extern unsigned int wimax_2016_0_75A_X;
extern int16_t wimax_2016_0_75A[504][15];
#endif

#ifdef Z_84
extern unsigned int wimax_2016_0_75B_D;
extern unsigned int wimax_2016_0_75B_X;
extern int16_t wimax_2016_0_75B[504][15];
#endif

#ifdef Z_84
extern unsigned int wimax_2016_0_83_D;
extern unsigned int wimax_2016_0_83_X;
extern int16_t wimax_2016_0_83[336][20];
#endif

#ifdef Z_88
extern unsigned int wimax_2112_0_5_D;
extern unsigned int wimax_2112_0_5_X;
extern int16_t wimax_2112_0_5[1056][7];
#endif

#ifdef Z_88
extern unsigned int wimax_2112_0_66A_D;
extern unsigned int wimax_2112_0_66A_X;
extern int16_t wimax_2112_0_66A[704][10];
#endif

#ifdef Z_88
extern unsigned int wimax_2112_0_66B_D;
extern unsigned int wimax_2112_0_66B_X;
extern int16_t wimax_2112_0_66B[704][11];
#endif

#ifdef Z_88
extern unsigned int wimax_2112_0_75A_D;
extern unsigned int wimax_2112_0_75A_X;
extern int16_t wimax_2112_0_75A[528][15];
#endif

#ifdef Z_88
extern unsigned int wimax_2112_0_75B_D;
extern unsigned int wimax_2112_0_75B_X;
extern int16_t wimax_2112_0_75B[528][15];
#endif

#ifdef Z_88
extern unsigned int wimax_2112_0_83_D;
extern unsigned int wimax_2112_0_83_X;
extern int16_t wimax_2112_0_83[352][20];
#endif

#ifdef Z_92
extern unsigned int wimax_2208_0_5_D;
extern unsigned int wimax_2208_0_5_X;
extern int16_t wimax_2208_0_5[1104][7];
#endif

#ifdef Z_92
extern unsigned int wimax_2208_0_66A_D;
extern unsigned int wimax_2208_0_66A_X;
extern int16_t wimax_2208_0_66A[736][10];
#endif

#ifdef Z_92
extern unsigned int wimax_2208_0_66B_D;
extern unsigned int wimax_2208_0_66B_X;
extern int16_t wimax_2208_0_66B[736][11];
#endif

#ifdef Z_92
extern unsigned int wimax_2208_0_75A_D;
extern unsigned int wimax_2208_0_75A_X;
extern int16_t wimax_2208_0_75A[552][15];
#endif

#ifdef Z_92
extern unsigned int wimax_2208_0_75B_D;
extern unsigned int wimax_2208_0_75B_X;
extern int16_t wimax_2208_0_75B[552][15];
#endif

#ifdef Z_92
extern unsigned int wimax_2208_0_83_D;
extern unsigned int wimax_2208_0_83_X;
extern int16_t wimax_2208_0_83[368][20];
#endif

#ifdef Z_96
extern unsigned int wimax_2304_0_5_D;
extern unsigned int wimax_2304_0_5_X;
extern int16_t wimax_2304_0_5[1152][7];
#endif

#ifdef Z_96
extern unsigned int wimax_2304_0_66A_D;
extern unsigned int wimax_2304_0_66A_X;
extern int16_t wimax_2304_0_66A[768][10];
#endif

#ifdef Z_96
extern unsigned int wimax_2304_0_66B_D;
extern unsigned int wimax_2304_0_66B_X;
extern int16_t wimax_2304_0_66B[768][11];
#endif

#ifdef Z_96
extern unsigned int wimax_2304_0_75A_D;
extern unsigned int wimax_2304_0_75A_X;
extern int16_t wimax_2304_0_75A[576][15];
#endif

#ifdef Z_96
extern unsigned int wimax_2304_0_75B_D;
extern unsigned int wimax_2304_0_75B_X;
extern int16_t wimax_2304_0_75B[576][15];
#endif

#ifdef Z_96
extern unsigned int wimax_2304_0_83_D;
extern unsigned int wimax_2304_0_83_X;
extern int16_t wimax_2304_0_83[384][20];
#endif

#endif  // MATRICES_H
