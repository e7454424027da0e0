// Generated by gen_constants.py -- do not edit.
// BLS12-381 constants, 14 x 28-bit limbs (little-endian), internal
// Montgomery radix R' = 2^392.  'mont' values are canonical v*R' mod p.
#pragma once
#include <stdint.h>
namespace bls {
// p
constexpr int32_t P_L[14] = {0xfffaaab, 0xfefffff, 0x3ffffb9, 0xfffeb15, 0x6241eab, 0xa0f6b0f, 0xf6730d2, 0xf38512b, 0x4774b84, 0x4bacd76, 0xba7b643, 0xe69a4b1, 0x1ea397f, 0x001a011};
constexpr uint32_t NPRIME = 0xffcfffd;  // -p^-1 mod 2^28
// 1 in internal Montgomery form
constexpr int32_t ONE_M[14] = {0x347fcb8, 0xd800000, 0x002b119, 0x0cde6d2, 0xc7212e0, 0x83a2090, 0x037669f, 0xda0f73e, 0x9b09b42, 0x1297bb0, 0x515d98f, 0x012ca7c, 0x659fcfa, 0x000577a};
// curve constant b = 4 (G1)
constexpr int32_t B_G1_M[14] = {0xd1ff2e0, 0x6000000, 0x00ac467, 0x3379b48, 0x1c84b80, 0x0e88243, 0x0dd9a7e, 0x683dcf8, 0x6c26d0b, 0x4a5eec2, 0x457663c, 0x04b29f1, 0x967f3e8, 0x0015de9};
// 2^400 mod p: mont(x, C_TO_INT) maps x*2^384 -> x*2^392
constexpr int32_t C_TO_INT[14] = {0x80e6299, 0x3500034, 0xeb12856, 0xdeb2699, 0xc988670, 0x4ef6697, 0x70983e8, 0xa4e6fe9, 0x3e8a053, 0xecf271e, 0xc20d323, 0x6eb6385, 0x47f1286, 0x00156da};
// 2^384 mod p: mont(x, C_TO_REF) maps x*2^392 -> x*2^384
constexpr int32_t C_TO_REF[14] = {0x002fffd, 0x0900000, 0xc000276, 0x000bc40, 0x8baebf4, 0x5753c75, 0x55f4898, 0x7052574, 0x7ce5853, 0x56ec6d7, 0x71a97a2, 0xe4935c0, 0xec3fa80, 0x0015f65};
// R'^2 mod p: mont(v, C_R2_INT) maps plain v -> v*R'
constexpr int32_t C_R2_INT[14] = {0x10370ed, 0x6d1c345, 0xe243d62, 0xec45c53, 0x3b1d65a, 0x093317d, 0xb4f36a0, 0x5d74088, 0xc10ea72, 0x865d118, 0x7320a75, 0xfd5cd50, 0xcc8a759, 0x000c8d4};
// R'^3 mod p: mont(x^-1, C_R3_INT) maps the inverse of x = aR' to a^-1 R'
constexpr int32_t C_R3_INT[14] = {0x1f7b890, 0x294cc4d, 0x9f3af22, 0xb5ba56c, 0xcb5c0cc, 0xc0d975c, 0xc89a8c5, 0x6c968b4, 0x22672ea, 0x91de8c9, 0x35652a6, 0x84977c8, 0x424bbb9, 0x00141ab};
// (1+u)^((p-1)/3) = (0, this)
constexpr int32_t FROB6_C1_1[14] = {0x2421b59, 0xbee4867, 0x1d31002, 0x4760184, 0x4cc5086, 0xc76dc00, 0xaae891b, 0xac70ad2, 0xfe377c4, 0xe4686b8, 0x5ed1568, 0x8f5a180, 0x02b5c1f, 0x000d1a4};
// (1+u)^(2(p-1)/3) = (this, 0)
constexpr int32_t FROB6_C2_0[14] = {0x58a1811, 0x96e4867, 0x1d5c11c, 0x543e856, 0x13e6366, 0x4b0fc91, 0xae5efbb, 0x8680210, 0x9941307, 0xf700269, 0xb02eef7, 0x9086bfc, 0x6855919, 0x001291e};
// (1+u)^((p-1)/6) = (this, .)
constexpr int32_t FROB12_0[14] = {0x9f1ba38, 0xed52b31, 0x3131f18, 0x932815a, 0x35bde3f, 0x7c4a4df, 0xc6f7465, 0x266b7cc, 0x2acd4ff, 0xcae398d, 0x243b688, 0xa613121, 0x72d376f, 0x0001c3e};
// (1+u)^((p-1)/6) = (., this)
constexpr int32_t FROB12_1[14] = {0x60df073, 0x11ad4ce, 0x0ece0a1, 0x6cd69bb, 0x2c8406c, 0x24ac630, 0x2f7bc6d, 0xcd1995f, 0x1ca7685, 0x80c93e9, 0x963ffba, 0x4087390, 0xabd0210, 0x00183d2};
// cube root of unity: (beta x, y) = [-x^2](x, y) on G1
constexpr int32_t BETA_M[14] = {0xa75929a, 0x681b798, 0x22a3e9d, 0xabc02bf, 0x4e5bb45, 0x55e6e7e, 0x4814117, 0x6d04f1b, 0xae3387d, 0x54acb0c, 0x0a4c74b, 0x56138b5, 0xb64e066, 0x00076f2};
// G2 curve constant b' = 4(1+u) = (this, this)
constexpr int32_t B_G2_M[14] = {0xd1ff2e0, 0x6000000, 0x00ac467, 0x3379b48, 0x1c84b80, 0x0e88243, 0x0dd9a7e, 0x683dcf8, 0x6c26d0b, 0x4a5eec2, 0x457663c, 0x04b29f1, 0x967f3e8, 0x0015de9};
// psi: cx.c0
constexpr int32_t PSI_X0[14] = {0x0000000, 0x0000000, 0x0000000, 0x0000000, 0x0000000, 0x0000000, 0x0000000, 0x0000000, 0x0000000, 0x0000000, 0x0000000, 0x0000000, 0x0000000, 0x0000000};
// psi: cx.c1
constexpr int32_t PSI_X1[14] = {0x58a1811, 0x96e4867, 0x1d5c11c, 0x543e856, 0x13e6366, 0x4b0fc91, 0xae5efbb, 0x8680210, 0x9941307, 0xf700269, 0xb02eef7, 0x9086bfc, 0x6855919, 0x001291e};
// psi: cy.c0
constexpr int32_t PSI_Y0[14] = {0xcc17b84, 0xcc5da55, 0x1835de7, 0x3e1e677, 0x9e4ae31, 0x9b9a07a, 0xd662557, 0xb7f1997, 0x71cc4da, 0xa667f92, 0x65115fe, 0x4a3370c, 0xe5b746a, 0x000d16d};
// psi: cy.c1
constexpr int32_t PSI_Y1[14] = {0x33e2f27, 0x32a25aa, 0x27ca1d2, 0xc1e049e, 0xc3f707a, 0x055ca94, 0x2010b7b, 0x3b93794, 0xd5a86aa, 0xa544de3, 0x556a044, 0x9c66da5, 0x38ec515, 0x000cea3};
// (p+1)/4 for the square root, p-2 for the Fermat inverse, (p-1)/2 for Euler's criterion
constexpr uint32_t SQRT_EXP[12] = {0xffffeaabu, 0xee7fbfffu, 0xac54ffffu, 0x07aaffffu, 0x3dac3d89u, 0xd9cc34a8u, 0x3ce144afu, 0xd91dd2e1u, 0x90d2eb35u, 0x92c6e9edu, 0x8e5ff9a6u, 0x0680447au};
constexpr uint32_t INV_EXP[12] = {0xffffaaa9u, 0xb9feffffu, 0xb153ffffu, 0x1eabfffeu, 0xf6b0f624u, 0x6730d2a0u, 0xf38512bfu, 0x64774b84u, 0x434bacd7u, 0x4b1ba7b6u, 0x397fe69au, 0x1a0111eau};
constexpr uint32_t HALF_P32[12] = {0xffffd555u, 0xdcff7fffu, 0x58a9ffffu, 0x0f55ffffu, 0x7b587b12u, 0xb3986950u, 0x79c2895fu, 0xb23ba5c2u, 0x21a5d66bu, 0x258dd3dbu, 0x1cbff34du, 0x0d0088f5u};
// (p-3)/4 for the square root in Fp2
constexpr uint32_t P34_EXP[12] = {0xffffeaaau, 0xee7fbfffu, 0xac54ffffu, 0x07aaffffu, 0x3dac3d89u, 0xd9cc34a8u, 0x3ce144afu, 0xd91dd2e1u, 0x90d2eb35u, 0x92c6e9edu, 0x8e5ff9a6u, 0x0680447au};
constexpr uint32_t P32[12] = {0xffffaaabu, 0xb9feffffu, 0xb153ffffu, 0x1eabfffeu, 0xf6b0f624u, 0x6730d2a0u, 0xf38512bfu, 0x64774b84u, 0x434bacd7u, 0x4b1ba7b6u, 0x397fe69au, 0x1a0111eau};
constexpr uint32_t R_ORDER32[8] = {0x00000001u, 0xffffffffu, 0xfffe5bfeu, 0x53bda402u, 0x09a1d805u, 0x3339d808u, 0x299d7d48u, 0x73eda753u};
constexpr uint64_t BLS_X = 0xd201000000010000ull;  // x = -BLS_X
// (p+1)/4 as a 4-bit sliding window: {squarings, odd digit} (see gen_constants.py)
constexpr int SQRT_WIN_LEN = 79;
constexpr uint8_t SQRT_WIN[79][2] = {{4, 13}, {9, 1}, {4, 1}, {7, 15}, {4, 5}, {6, 7}, {6, 11}, {4, 15}, {4, 15}, {6, 13}, {6, 13}, {6, 9}, {3, 3}, {7, 13}, {4, 13}, {6, 15}, {5, 13}, {4, 9}, {8, 13}, {6, 11}, {3, 5}, {3, 3}, {6, 13}, {4, 7}, {3, 3}, {3, 1}, {6, 7}, {4, 7}, {5, 9}, {4, 7}, {8, 9}, {3, 7}, {5, 7}, {7, 5}, {7, 9}, {5, 11}, {4, 15}, {3, 3}, {5, 7}, {4, 3}, {8, 13}, {5, 5}, {2, 1}, {9, 15}, {5, 13}, {3, 3}, {8, 15}, {3, 3}, {7, 9}, {9, 15}, {4, 5}, {5, 11}, {4, 15}, {4, 15}, {4, 15}, {3, 7}, {5, 11}, {6, 5}, {5, 9}, {4, 15}, {4, 15}, {4, 15}, {4, 15}, {4, 13}, {2, 3}, {6, 15}, {4, 15}, {5, 15}, {4, 15}, {4, 15}, {4, 15}, {4, 15}, {4, 15}, {4, 15}, {4, 15}, {3, 5}, {4, 5}, {4, 5}, {3, 3}};
}  // namespace bls
