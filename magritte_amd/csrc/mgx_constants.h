/*
 * Physical and numerical constants of the formal solver, with the values and
 * derived-constant expressions of Magritte's src/tools/constants.hpp:15-40
 * (CC_SQUARED is the literal 8.98755179E+16 there, not CC*CC).
 */
#pragma once

#define MGX_PI 3.14159265358979323846
#define MGX_2_SQRTPI 1.12837916709551257390 /* 2/sqrt(pi), M_2_SQRTPI */
#define MGX_INVERSE_SQRT_PI (0.5 * MGX_2_SQRTPI)
#define MGX_FOUR_PI (4.0 * MGX_PI)

#define MGX_CC 2.99792458E+8
#define MGX_HH 6.62607004E-34
#define MGX_KB 1.38064852E-23
#define MGX_AMU 1.66053904E-27
#define MGX_T_CMB 2.7254800
#define MGX_CC_SQUARED 8.98755179E+16

#define MGX_TWO_KB_OVER_AMU_CC_SQUARED (2.0 * MGX_KB / (MGX_AMU * MGX_CC_SQUARED))
#define MGX_TWO_HH_OVER_CC_SQUARED (2.0 * MGX_HH / MGX_CC_SQUARED)
#define MGX_HH_OVER_KB (MGX_HH / MGX_KB)
#define MGX_HH_OVER_FOUR_PI (MGX_HH / MGX_FOUR_PI)
