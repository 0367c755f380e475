// Plane table: every device array is plane p of one allocation,
// base + p * layout.total (doubles) or bytes + p * layout.total (flags).
#pragma once

#include "gcs_layout.hpp"

namespace gcs {

// ---- static geometry (computed once at setup) ----------------------------
// Cell: volume, centroid, min CFL length, and the calibrated-recovery
// elimination constants (see recover_cell in gcs_math.hpp).
constexpr int P_VOL = 0;
constexpr int P_CEN = 1;       // +3
constexpr int P_CFL = 4;       // +3 CFL length per direction
constexpr int P_REC = 7;       // +9: inv0 f01 f02 inv1 f10 f12 inv2 f20 f21
// Faces, direction d: P_FACE + 16 d + {area, area vector(3), centroid(3),
// normal(3), tangent1(3), tangent2(3)} of the calibrated face frame.
constexpr int P_FACE = 16;
constexpr int FO_AREA = 0, FO_AV = 1, FO_CEN = 4, FO_N = 7, FO_T1 = 10,
              FO_T2 = 13, FACE_PLANES = 16;
// Edges, direction e: P_EDGE + 15 e + {length, normal(3), tangent1(3),
// tangent2(3), transverse solve n1x n1y n2x n2y 1/det (5)}.
constexpr int P_EDGE = P_FACE + 3 * FACE_PLANES;  // 64
constexpr int EO_LEN = 0, EO_N = 1, EO_T1 = 4, EO_T2 = 7, EO_SOLVE = 10,
              EDGE_PLANES = 15;
// Background (split-field B0), static.
constexpr int P_B0C = P_EDGE + 3 * EDGE_PLANES;   // 109, +3 cell B0
constexpr int P_FORCE = P_B0C + 3;                // +3 cell dpB0 force
constexpr int P_B0F = P_FORCE + 3;                // +9 face B0, 3 per dir
constexpr int P_B0FLUX = P_B0F + 9;               // +3 face B0 flux
constexpr int P_B0N = P_B0FLUX + 3;               // +3 face B0 flux / area
constexpr int P_B0E = P_B0N + 3;                  // +9 edge B0, 3 per dir
constexpr int P_B0ET = P_B0E + 9;                 // +6 edge B0 . (t1, t2)
constexpr int P_VERT = P_B0ET + 6;                // +3 vertex coordinates
constexpr int STATIC_PLANES = P_VERT + 3;         // 148

// ---- dynamic state --------------------------------------------------------
// Two time levels of U (5), cell B (3) and face flux (3) alternate roles
// every step (level s is current, 1 - s is old).
constexpr int P_U = STATIC_PLANES;                // +2*5
constexpr int P_B = P_U + 10;                     // +2*3
constexpr int P_PHI = P_B + 6;                    // +2*3
constexpr int P_UPRED = P_PHI + 6;                // +5
constexpr int P_BPRED = P_UPRED + 5;              // +3
constexpr int P_PHIPRED = P_BPRED + 3;            // +3
constexpr int P_FLUID = P_PHIPRED + 3;            // +15 (5 per dir)
constexpr int P_MAXWELL = P_FLUID + 15;           // +9 (3 per dir)
constexpr int P_EMF = P_MAXWELL + 9;              // +3
constexpr int P_FVEL = P_EMF + 3;                 // +9 face velocity, 3 per dir
constexpr int P_WPRED = P_FVEL + 9;                // +5 rho v p of the predicted state
                                                   // (sweep floors; NaN: no conversion)
constexpr int P_WV = P_WPRED + 5;                  // +3 vol * m / max(rho, EMF floor)
constexpr int P_BN = P_WV + 3;                     // +3 Phi* / area per direction
constexpr int PLANES = P_BN + 3;
// Setup reuses the flux planes as scratch before the first step.
constexpr int P_SCRATCH_STRESS = P_FLUID;         // +9 background face stress
constexpr int P_SCRATCH_AREA = P_FLUID + 9;       // +3 uncalibrated face area
constexpr int P_SCRATCH_LINE = P_FLUID + 12;      // +3 edge line integrals of A

// ---- FP32 planes of the mixed form ----------------------------------------
// Problem::planes32 + q * layout.total (floats), allocated when the mixed
// form is first selected. The FP32 kernels read their inputs and write
// their results here at the precision they compute in, so the predictor's
// caches, the face and edge geometry they use, the face velocities and the
// face fluxes cross HBM at 4 bytes instead of 8.
constexpr int M_VOL = 0;                           // cell volume
constexpr int M_WPRED = M_VOL + 1;                 // +5 (as P_WPRED)
constexpr int M_BPRED = M_WPRED + 5;               // +3
constexpr int M_UPRED = M_BPRED + 3;               // +5
constexpr int M_WV = M_UPRED + 5;                  // +3
constexpr int M_BN = M_WV + 3;                     // +3
// Faces, direction d: M_FACE + 14 d + {area, normal(3), tangent1(3),
// tangent2(3), B0(3), B0 flux / area}.
constexpr int M_FACE = M_BN + 3;
constexpr int MF_AREA = 0, MF_N = 1, MF_T1 = 4, MF_T2 = 7, MF_B0 = 10, MF_B0N = 13,
              M_FACE_PLANES = 14;
constexpr int M_FVEL = M_FACE + 3 * M_FACE_PLANES; // +9 face velocity, 3 per dir
constexpr int M_FLUID = M_FVEL + 9;                // +15 area-scaled, 5 per dir
constexpr int M_MAXWELL = M_FLUID + 15;            // +9, 3 per dir
// Edges, direction e: M_EDGE + 14 e + {length, tangent1(3), tangent2(3),
// transverse solve(5), B0 . (t1, t2)}.
constexpr int M_EDGE = M_MAXWELL + 9;
constexpr int ME_LEN = 0, ME_T1 = 1, ME_T2 = 4, ME_SOLVE = 7, ME_B0T = 12,
              M_EDGE_PLANES = 14;
constexpr int PLANES32 = M_EDGE + 3 * M_EDGE_PLANES;  // 137

// Byte planes.
constexpr int Q_VALID = 0;   // +3 edge validity (nonorthogonal_grid.c:391-403)
constexpr int Q_EBOK = 3;    // +3 edge transverse solve well posed
constexpr int Q_PIV = 6;     // cell recovery pivots: piv0 | piv1 << 2, 0xff bad
constexpr int BYTE_PLANES = 7;

inline __host__ __device__ int plane_u(int level, int q) { return P_U + 5 * level + q; }
inline __host__ __device__ int plane_b(int level, int c) { return P_B + 3 * level + c; }
inline __host__ __device__ int plane_phi(int level, int d) { return P_PHI + 3 * level + d; }

}  // namespace gcs
