// Generated by gen_dop853_rkn.py from the tableau in dop853_dev.h -- do not edit.
// DOP853 in Nystrom form for y = (q, p), y' = (p, a): see the generator's
// docstring for the derivation.  A[j] is the acceleration of stage j
// (A[13] = a(q_new), the FSAL stage); q0, p0 the state at the start of the
// step, pn the new velocity.
#pragma once
#include "integrators_dev.h"

namespace ghrk {
constexpr double qp2 = 0x1.aee6838dae63ap-5;
constexpr double qp3 = 0x1.432ce2aa42cacp-4;
constexpr double qa3_1 = 0x1.97fa4bd857c3ap-9;
constexpr double qp4 = 0x1.e4c353ff64302p-4;
constexpr double qa4_1 = 0x1.caf9955362bc1p-10;
constexpr double qa4_2 = 0x1.583b2ffe8a0d1p-8;
constexpr double qp5 = 0x1.2068c499c08d9p-2;
constexpr double qa5_1 = 0x1.44ebe53a4526dp-7;
constexpr double qa5_2 = -0x1.accbfefe6d3eep-5;
constexpr double qa5_3 = 0x1.503e7575108dfp-4;
constexpr double qp6 = 0x1.5555555555555p-2;
constexpr double qa6_1 = 0x1.217d036868b4bp-5;
constexpr double qa6_3 = -0x1.8879abc79e49bp-4;
constexpr double qa6_4 = 0x1.db4962f6f8283p-4;
constexpr double qp7 = 0x1.0000000000000p-2;
constexpr double qa7_1 = 0x1.35fe4221648cap-6;
constexpr double qa7_3 = -0x1.388b7e0e7b3ffp-5;
constexpr double qa7_4 = 0x1.af9d9aa51cc97p-5;
constexpr double qa7_5 = -0x1.2113da753cfcdp-9;
constexpr double qp8 = 0x1.3b13b13b13b13p-2;
constexpr double qa8_1 = 0x1.f68031439116bp-6;
constexpr double qa8_3 = -0x1.46acff29807b4p-4;
constexpr double qa8_4 = 0x1.915f39fadea92p-4;
constexpr double qa8_5 = -0x1.7542aa816a524p-10;
constexpr double qa8_6 = -0x1.3101454da6cdfp-13;
constexpr double qp9 = 0x1.4d74d74d74d72p-1;
constexpr double qa9_1 = -0x1.37e44a3a3d283p-3;
constexpr double qa9_3 = 0x1.e0eec7f1761d7p-2;
constexpr double qa9_4 = -0x1.171b710b3b63dp-4;
constexpr double qa9_5 = 0x1.5f0192e399b75p-7;
constexpr double qa9_6 = 0x1.3f75068de373dp-2;
constexpr double qa9_7 = -0x1.707642fcde52cp-2;
constexpr double qp10 = 0x1.3333333333323p-1;
constexpr double qa10_1 = -0x1.c36897302cb72p-4;
constexpr double qa10_3 = 0x1.34853e327dabcp-2;
constexpr double qa10_4 = 0x1.422e368da0519p-4;
constexpr double qa10_5 = 0x1.f9439ccc790afp-6;
constexpr double qa10_6 = -0x1.47464f4598606p-2;
constexpr double qa10_7 = -0x1.5ecf653d1571ep-1;
constexpr double qa10_8 = 0x1.c4b612b71fecap-1;
constexpr double qp11 = 0x1.b6db6db6db6e3p-1;
constexpr double qa11_1 = 0x1.7d1f3e8130ef8p-2;
constexpr double qa11_3 = -0x1.029902a47a1c0p-1;
constexpr double qa11_4 = -0x1.d893209d33c95p-2;
constexpr double qa11_5 = -0x1.0aff864b4e23dp-4;
constexpr double qa11_6 = 0x1.064642de96f8ap+2;
constexpr double qa11_7 = 0x1.f233f254578b8p+1;
constexpr double qa11_8 = -0x1.c19e2850052fdp+2;
constexpr double qa11_9 = 0x1.fb72c96a00a97p-5;
constexpr double qp12 = 0x1.0000000000003p+0;
constexpr double qa12_1 = -0x1.87b7e9647bf2fp-1;
constexpr double qa12_3 = 0x1.ab6ad6747361cp-1;
constexpr double qa12_4 = 0x1.c1858e30e981ep+0;
constexpr double qa12_5 = 0x1.dc3c5a4906309p-3;
constexpr double qa12_6 = 0x1.7ceb44dbdbab7p+3;
constexpr double qa12_7 = -0x1.e74b71c5bf258p+0;
constexpr double qa12_8 = -0x1.5e707259e25a6p+3;
constexpr double qa12_9 = 0x1.5a624e7cc0091p+0;
constexpr double qa12_10 = -0x1.f5d400c8d170ap+0;
constexpr double qp14 = 0x1.bb96af038e29fp-4;
constexpr double qn14 = -0x1.0fe8ab4fa4830p-7;
constexpr double qa14_1 = 0x1.68c49573440a6p-8;
constexpr double qa14_4 = 0x1.7e6a7f2ad256bp-54;
constexpr double qa14_5 = 0x1.04b0a1a7115a4p-56;
constexpr double qa14_6 = -0x1.80c2589368fc0p-2;
constexpr double qa14_7 = -0x1.a6a8a994c5089p-4;
constexpr double qa14_8 = 0x1.da3a08deaad3bp-2;
constexpr double qa14_9 = -0x1.98163c4249041p-5;
constexpr double qa14_10 = 0x1.18cf0ca21fb64p-4;
constexpr double qa14_11 = 0x1.3f1a74130d863p-8;
constexpr double qp15 = 0x1.9a4c1a0a38d24p-3;
constexpr double qn15 = -0x1.6500e13e7149bp-12;
constexpr double qa15_1 = 0x1.449bc95143326p-7;
constexpr double qa15_4 = 0x1.be05263b64357p-61;
constexpr double qa15_5 = -0x1.7e31c5d4db024p-62;
constexpr double qa15_6 = -0x1.8ef067bf3cc87p-8;
constexpr double qa15_7 = 0x1.89c4857e353efp-5;
constexpr double qa15_8 = -0x1.3a3287c6b66abp-5;
constexpr double qa15_9 = -0x1.5b93ed18268e3p-6;
constexpr double qa15_10 = 0x1.b5cbc22686888p-6;
constexpr double qa15_11 = 0x1.7053977170015p-10;
constexpr double qa15_12 = 0x1.1858c6dcd84e9p-10;
constexpr double qa15_13 = -0x1.336491b39821dp-10;
constexpr double qp16 = 0x1.8ef043044b278p-1;
constexpr double qn16 = -0x1.6ebeec24871d4p-10;
constexpr double qa16_1 = 0x1.6f8e91dffea72p-2;
constexpr double qa16_4 = 0x1.5b9bd25fe7076p-54;
constexpr double qa16_5 = 0x1.c0d0bd6e57093p-55;
constexpr double qa16_6 = 0x1.2c5e04cccd6a2p+3;
constexpr double qa16_7 = 0x1.debef435f73f4p+2;
constexpr double qa16_8 = -0x1.f7973e4f63f9cp+3;
constexpr double qa16_9 = -0x1.76d766211458fp-2;
constexpr double qa16_10 = 0x1.ceac7b7346c3cp-2;
constexpr double qa16_11 = 0x1.9c4a53328965fp-6;
constexpr double qa16_12 = 0x1.341c6537f4956p-6;
constexpr double qa16_13 = -0x1.5dae98afd0f7ap-6;
constexpr double qa16_14 = -0x1.4b0b814d442a4p+0;
constexpr double bp = 0x1.0000000000000p+0;
constexpr double ba_1 = 0x1.bcc6368d1177dp-5;
constexpr double ba_4 = 0x1.453149357e02cp-54;
constexpr double ba_5 = -0x1.42c3b0243cc92p-58;
constexpr double ba_6 = 0x1.7bc2918e723d9p+1;
constexpr double ba_7 = 0x1.6b2be346040b9p+0;
constexpr double ba_8 = -0x1.0109b7c05dbcdp+2;
constexpr double ba_9 = 0x1.bc738374d6d62p-4;
constexpr double ba_10 = -0x1.f299db205e66fp-5;
constexpr double ba_11 = 0x1.d74f62f6d9a7dp-6;
constexpr double e3p = 0x1.d000000000000p-54;
constexpr double e3a_1 = -0x1.d0bf40d812c54p-2;
constexpr double e3a_4 = 0x1.5970eb2982213p+1;
constexpr double e3a_5 = 0x1.5cd04f410d934p-1;
constexpr double e3a_6 = -0x1.0e2a5d3ab05f5p+4;
constexpr double e3a_7 = -0x1.bf9cafd65a23fp+3;
constexpr double e3a_8 = 0x1.bf6359ba252c3p+4;
constexpr double e3a_9 = 0x1.3788f4dcb4394p-2;
constexpr double e3a_10 = -0x1.55874d65582b8p-2;
constexpr double e3a_11 = 0x1.dd8f901379b8dp-7;
constexpr double e5p = -0x1.0000000000000p-58;
constexpr double e5a_1 = -0x1.825beab80ca50p-3;
constexpr double e5a_4 = 0x1.fe0fe8f118cc7p-1;
constexpr double e5a_5 = 0x1.e352b7cd95f77p-3;
constexpr double e5a_6 = -0x1.6d6489c76d7b7p+1;
constexpr double e5a_7 = -0x1.054cbd65f2e72p+2;
constexpr double e5a_8 = 0x1.8274300ebeaafp+2;
constexpr double e5a_9 = 0x1.95587b9df5488p-2;
constexpr double e5a_10 = -0x1.0d460aa2428f6p-1;
constexpr double e5a_11 = -0x1.d74f62f6d9a7dp-7;
constexpr double dp4 = 0x1.6c81218b7f008p-4;
constexpr double dn4 = -0x1.6c81218b7f07cp-4;
constexpr double da4_1 = 0x1.6ed92316b375fp+1;
constexpr double da4_4 = 0x1.0c9e86a35a0adp-49;
constexpr double da4_5 = 0x1.1f747d86e34fdp-52;
constexpr double da4_6 = 0x1.ee87f2f35db26p+4;
constexpr double da4_7 = -0x1.871bed7a9871fp+4;
constexpr double da4_8 = -0x1.1f7ca44ec0e92p+5;
constexpr double da4_9 = -0x1.eafa722f551ccp+1;
constexpr double da4_10 = 0x1.e19d00616623cp+1;
constexpr double da4_11 = 0x1.1cc472176cfe2p-1;
constexpr double da4_12 = 0x1.1214a4134bb52p-3;
constexpr double da4_13 = -0x1.214cc77f1b19ap-3;
constexpr double da4_14 = -0x1.cbfcb6af597e2p+3;
constexpr double da4_15 = 0x1.44c07b157c3b9p+5;
constexpr double dp5 = -0x1.f64fc65250f80p+3;
constexpr double dn5 = 0x1.f64fc65250f7cp+3;
constexpr double da5_1 = -0x1.2c6024278f90bp+4;
constexpr double da5_4 = 0x1.b41d1226be642p-48;
constexpr double da5_5 = 0x1.152b6178b6f1ep-49;
constexpr double da5_6 = -0x1.842adab6a43c0p+7;
constexpr double da5_7 = 0x1.e6e48850ccc59p+7;
constexpr double da5_8 = 0x1.76c08b0b6ab2cp+7;
constexpr double da5_9 = 0x1.6ccea399ad33bp+4;
constexpr double da5_10 = -0x1.aaac90503517ep+4;
constexpr double da5_11 = -0x1.908cb2fc6b6e4p+2;
constexpr double da5_12 = -0x1.e9f5997d4cfffp-3;
constexpr double da5_13 = 0x1.b1176ffc70093p-3;
constexpr double da5_14 = 0x1.a0fef6065b0fep+6;
constexpr double da5_15 = -0x1.47c22867efa94p+8;
constexpr double dp6 = 0x1.639c3efff567ep+1;
constexpr double dn6 = -0x1.639c3efff56d2p+1;
constexpr double da6_1 = -0x1.1b27815c1179cp+3;
constexpr double da6_4 = -0x1.b5836caf5c1c6p-47;
constexpr double da6_5 = -0x1.0c80fe01d155cp-48;
constexpr double da6_6 = -0x1.db8fc0883544dp+7;
constexpr double da6_7 = -0x1.0b89c821c341dp+1;
constexpr double da6_8 = 0x1.344b8d01d0464p+8;
constexpr double da6_9 = 0x1.1bf22e687c62ap+1;
constexpr double da6_10 = 0x1.b78b13e5e2785p+1;
constexpr double da6_11 = -0x1.3f6327ac86a73p-9;
constexpr double da6_12 = -0x1.b176927adeccfp-2;
constexpr double da6_13 = 0x1.d0ec54463c8cap-2;
constexpr double da6_14 = 0x1.7a1a943a2be63p+5;
constexpr double da6_15 = -0x1.b6f6bdbd4b894p+6;
constexpr double dp7 = 0x1.5c4484e37f77cp+5;
constexpr double dn7 = -0x1.5c4484e37f77ep+5;
constexpr double da7_1 = 0x1.381ad4ad41f89p+6;
constexpr double da7_4 = 0x1.b2555d074ff61p-47;
constexpr double da7_5 = -0x1.aa1ec87387c7ap-48;
constexpr double da7_6 = 0x1.12a3ccf5452a6p+10;
constexpr double da7_7 = -0x1.c6735f00f589cp+9;
constexpr double da7_8 = -0x1.230d8043a414cp+10;
constexpr double da7_9 = -0x1.173115b0116fap+6;
constexpr double da7_10 = 0x1.09c4334879105p+6;
constexpr double da7_11 = 0x1.3fe44b350de79p+4;
constexpr double da7_12 = 0x1.6d8f9c0ab2a63p-1;
constexpr double da7_13 = -0x1.2729dac42ab0dp-1;
constexpr double da7_14 = -0x1.bedbb97825c75p+8;
constexpr double da7_15 = 0x1.5689384aa77aap+10;
}  // namespace ghrk

// position argument of stage 2
__device__ __forceinline__ GhVec3 gh_rkn_q2(double h, const GhVec3& q0, const GhVec3& p0,
                                             const GhVec3& pn, const GhVec3* A) {
  (void)pn;
  GhVec3 Q;
  Q.x = q0.x + h * (ghrk::qp2 * p0.x);
  Q.y = q0.y + h * (ghrk::qp2 * p0.y);
  Q.z = q0.z + h * (ghrk::qp2 * p0.z);
  return Q;
}

// position argument of stage 3
__device__ __forceinline__ GhVec3 gh_rkn_q3(double h, const GhVec3& q0, const GhVec3& p0,
                                             const GhVec3& pn, const GhVec3* A) {
  (void)pn;
  GhVec3 Q;
  Q.x = q0.x + h * (ghrk::qp3 * p0.x + h * (ghrk::qa3_1 * A[1].x));
  Q.y = q0.y + h * (ghrk::qp3 * p0.y + h * (ghrk::qa3_1 * A[1].y));
  Q.z = q0.z + h * (ghrk::qp3 * p0.z + h * (ghrk::qa3_1 * A[1].z));
  return Q;
}

// position argument of stage 4
__device__ __forceinline__ GhVec3 gh_rkn_q4(double h, const GhVec3& q0, const GhVec3& p0,
                                             const GhVec3& pn, const GhVec3* A) {
  (void)pn;
  GhVec3 Q;
  Q.x = q0.x + h * (ghrk::qp4 * p0.x + h * (ghrk::qa4_1 * A[1].x + ghrk::qa4_2 * A[2].x));
  Q.y = q0.y + h * (ghrk::qp4 * p0.y + h * (ghrk::qa4_1 * A[1].y + ghrk::qa4_2 * A[2].y));
  Q.z = q0.z + h * (ghrk::qp4 * p0.z + h * (ghrk::qa4_1 * A[1].z + ghrk::qa4_2 * A[2].z));
  return Q;
}

// position argument of stage 5
__device__ __forceinline__ GhVec3 gh_rkn_q5(double h, const GhVec3& q0, const GhVec3& p0,
                                             const GhVec3& pn, const GhVec3* A) {
  (void)pn;
  GhVec3 Q;
  Q.x = q0.x + h * (ghrk::qp5 * p0.x + h * (ghrk::qa5_1 * A[1].x + ghrk::qa5_2 * A[2].x + ghrk::qa5_3 * A[3].x));
  Q.y = q0.y + h * (ghrk::qp5 * p0.y + h * (ghrk::qa5_1 * A[1].y + ghrk::qa5_2 * A[2].y + ghrk::qa5_3 * A[3].y));
  Q.z = q0.z + h * (ghrk::qp5 * p0.z + h * (ghrk::qa5_1 * A[1].z + ghrk::qa5_2 * A[2].z + ghrk::qa5_3 * A[3].z));
  return Q;
}

// position argument of stage 6
__device__ __forceinline__ GhVec3 gh_rkn_q6(double h, const GhVec3& q0, const GhVec3& p0,
                                             const GhVec3& pn, const GhVec3* A) {
  (void)pn;
  GhVec3 Q;
  Q.x = q0.x + h * (ghrk::qp6 * p0.x + h * (ghrk::qa6_1 * A[1].x + ghrk::qa6_3 * A[3].x + ghrk::qa6_4 * A[4].x));
  Q.y = q0.y + h * (ghrk::qp6 * p0.y + h * (ghrk::qa6_1 * A[1].y + ghrk::qa6_3 * A[3].y + ghrk::qa6_4 * A[4].y));
  Q.z = q0.z + h * (ghrk::qp6 * p0.z + h * (ghrk::qa6_1 * A[1].z + ghrk::qa6_3 * A[3].z + ghrk::qa6_4 * A[4].z));
  return Q;
}

// position argument of stage 7
__device__ __forceinline__ GhVec3 gh_rkn_q7(double h, const GhVec3& q0, const GhVec3& p0,
                                             const GhVec3& pn, const GhVec3* A) {
  (void)pn;
  GhVec3 Q;
  Q.x = q0.x + h * (ghrk::qp7 * p0.x + h * (ghrk::qa7_1 * A[1].x + ghrk::qa7_3 * A[3].x + ghrk::qa7_4 * A[4].x + ghrk::qa7_5 * A[5].x));
  Q.y = q0.y + h * (ghrk::qp7 * p0.y + h * (ghrk::qa7_1 * A[1].y + ghrk::qa7_3 * A[3].y + ghrk::qa7_4 * A[4].y + ghrk::qa7_5 * A[5].y));
  Q.z = q0.z + h * (ghrk::qp7 * p0.z + h * (ghrk::qa7_1 * A[1].z + ghrk::qa7_3 * A[3].z + ghrk::qa7_4 * A[4].z + ghrk::qa7_5 * A[5].z));
  return Q;
}

// position argument of stage 8
__device__ __forceinline__ GhVec3 gh_rkn_q8(double h, const GhVec3& q0, const GhVec3& p0,
                                             const GhVec3& pn, const GhVec3* A) {
  (void)pn;
  GhVec3 Q;
  Q.x = q0.x + h * (ghrk::qp8 * p0.x + h * (ghrk::qa8_1 * A[1].x + ghrk::qa8_3 * A[3].x + ghrk::qa8_4 * A[4].x + ghrk::qa8_5 * A[5].x + ghrk::qa8_6 * A[6].x));
  Q.y = q0.y + h * (ghrk::qp8 * p0.y + h * (ghrk::qa8_1 * A[1].y + ghrk::qa8_3 * A[3].y + ghrk::qa8_4 * A[4].y + ghrk::qa8_5 * A[5].y + ghrk::qa8_6 * A[6].y));
  Q.z = q0.z + h * (ghrk::qp8 * p0.z + h * (ghrk::qa8_1 * A[1].z + ghrk::qa8_3 * A[3].z + ghrk::qa8_4 * A[4].z + ghrk::qa8_5 * A[5].z + ghrk::qa8_6 * A[6].z));
  return Q;
}

// position argument of stage 9
__device__ __forceinline__ GhVec3 gh_rkn_q9(double h, const GhVec3& q0, const GhVec3& p0,
                                             const GhVec3& pn, const GhVec3* A) {
  (void)pn;
  GhVec3 Q;
  Q.x = q0.x + h * (ghrk::qp9 * p0.x + h * (ghrk::qa9_1 * A[1].x + ghrk::qa9_3 * A[3].x + ghrk::qa9_4 * A[4].x + ghrk::qa9_5 * A[5].x + ghrk::qa9_6 * A[6].x + ghrk::qa9_7 * A[7].x));
  Q.y = q0.y + h * (ghrk::qp9 * p0.y + h * (ghrk::qa9_1 * A[1].y + ghrk::qa9_3 * A[3].y + ghrk::qa9_4 * A[4].y + ghrk::qa9_5 * A[5].y + ghrk::qa9_6 * A[6].y + ghrk::qa9_7 * A[7].y));
  Q.z = q0.z + h * (ghrk::qp9 * p0.z + h * (ghrk::qa9_1 * A[1].z + ghrk::qa9_3 * A[3].z + ghrk::qa9_4 * A[4].z + ghrk::qa9_5 * A[5].z + ghrk::qa9_6 * A[6].z + ghrk::qa9_7 * A[7].z));
  return Q;
}

// position argument of stage 10
__device__ __forceinline__ GhVec3 gh_rkn_q10(double h, const GhVec3& q0, const GhVec3& p0,
                                             const GhVec3& pn, const GhVec3* A) {
  (void)pn;
  GhVec3 Q;
  Q.x = q0.x + h * (ghrk::qp10 * p0.x + h * (ghrk::qa10_1 * A[1].x + ghrk::qa10_3 * A[3].x + ghrk::qa10_4 * A[4].x + ghrk::qa10_5 * A[5].x + ghrk::qa10_6 * A[6].x + ghrk::qa10_7 * A[7].x + ghrk::qa10_8 * A[8].x));
  Q.y = q0.y + h * (ghrk::qp10 * p0.y + h * (ghrk::qa10_1 * A[1].y + ghrk::qa10_3 * A[3].y + ghrk::qa10_4 * A[4].y + ghrk::qa10_5 * A[5].y + ghrk::qa10_6 * A[6].y + ghrk::qa10_7 * A[7].y + ghrk::qa10_8 * A[8].y));
  Q.z = q0.z + h * (ghrk::qp10 * p0.z + h * (ghrk::qa10_1 * A[1].z + ghrk::qa10_3 * A[3].z + ghrk::qa10_4 * A[4].z + ghrk::qa10_5 * A[5].z + ghrk::qa10_6 * A[6].z + ghrk::qa10_7 * A[7].z + ghrk::qa10_8 * A[8].z));
  return Q;
}

// position argument of stage 11
__device__ __forceinline__ GhVec3 gh_rkn_q11(double h, const GhVec3& q0, const GhVec3& p0,
                                             const GhVec3& pn, const GhVec3* A) {
  (void)pn;
  GhVec3 Q;
  Q.x = q0.x + h * (ghrk::qp11 * p0.x + h * (ghrk::qa11_1 * A[1].x + ghrk::qa11_3 * A[3].x + ghrk::qa11_4 * A[4].x + ghrk::qa11_5 * A[5].x + ghrk::qa11_6 * A[6].x + ghrk::qa11_7 * A[7].x + ghrk::qa11_8 * A[8].x + ghrk::qa11_9 * A[9].x));
  Q.y = q0.y + h * (ghrk::qp11 * p0.y + h * (ghrk::qa11_1 * A[1].y + ghrk::qa11_3 * A[3].y + ghrk::qa11_4 * A[4].y + ghrk::qa11_5 * A[5].y + ghrk::qa11_6 * A[6].y + ghrk::qa11_7 * A[7].y + ghrk::qa11_8 * A[8].y + ghrk::qa11_9 * A[9].y));
  Q.z = q0.z + h * (ghrk::qp11 * p0.z + h * (ghrk::qa11_1 * A[1].z + ghrk::qa11_3 * A[3].z + ghrk::qa11_4 * A[4].z + ghrk::qa11_5 * A[5].z + ghrk::qa11_6 * A[6].z + ghrk::qa11_7 * A[7].z + ghrk::qa11_8 * A[8].z + ghrk::qa11_9 * A[9].z));
  return Q;
}

// position argument of stage 12
__device__ __forceinline__ GhVec3 gh_rkn_q12(double h, const GhVec3& q0, const GhVec3& p0,
                                             const GhVec3& pn, const GhVec3* A) {
  (void)pn;
  GhVec3 Q;
  Q.x = q0.x + h * (ghrk::qp12 * p0.x + h * (ghrk::qa12_1 * A[1].x + ghrk::qa12_3 * A[3].x + ghrk::qa12_4 * A[4].x + ghrk::qa12_5 * A[5].x + ghrk::qa12_6 * A[6].x + ghrk::qa12_7 * A[7].x + ghrk::qa12_8 * A[8].x + ghrk::qa12_9 * A[9].x + ghrk::qa12_10 * A[10].x));
  Q.y = q0.y + h * (ghrk::qp12 * p0.y + h * (ghrk::qa12_1 * A[1].y + ghrk::qa12_3 * A[3].y + ghrk::qa12_4 * A[4].y + ghrk::qa12_5 * A[5].y + ghrk::qa12_6 * A[6].y + ghrk::qa12_7 * A[7].y + ghrk::qa12_8 * A[8].y + ghrk::qa12_9 * A[9].y + ghrk::qa12_10 * A[10].y));
  Q.z = q0.z + h * (ghrk::qp12 * p0.z + h * (ghrk::qa12_1 * A[1].z + ghrk::qa12_3 * A[3].z + ghrk::qa12_4 * A[4].z + ghrk::qa12_5 * A[5].z + ghrk::qa12_6 * A[6].z + ghrk::qa12_7 * A[7].z + ghrk::qa12_8 * A[8].z + ghrk::qa12_9 * A[9].z + ghrk::qa12_10 * A[10].z));
  return Q;
}

// position argument of stage 14
__device__ __forceinline__ GhVec3 gh_rkn_q14(double h, const GhVec3& q0, const GhVec3& p0,
                                             const GhVec3& pn, const GhVec3* A) {
  (void)pn;
  GhVec3 Q;
  Q.x = q0.x + h * (ghrk::qp14 * p0.x + ghrk::qn14 * pn.x + h * (ghrk::qa14_1 * A[1].x + ghrk::qa14_4 * A[4].x + ghrk::qa14_5 * A[5].x + ghrk::qa14_6 * A[6].x + ghrk::qa14_7 * A[7].x + ghrk::qa14_8 * A[8].x + ghrk::qa14_9 * A[9].x + ghrk::qa14_10 * A[10].x + ghrk::qa14_11 * A[11].x));
  Q.y = q0.y + h * (ghrk::qp14 * p0.y + ghrk::qn14 * pn.y + h * (ghrk::qa14_1 * A[1].y + ghrk::qa14_4 * A[4].y + ghrk::qa14_5 * A[5].y + ghrk::qa14_6 * A[6].y + ghrk::qa14_7 * A[7].y + ghrk::qa14_8 * A[8].y + ghrk::qa14_9 * A[9].y + ghrk::qa14_10 * A[10].y + ghrk::qa14_11 * A[11].y));
  Q.z = q0.z + h * (ghrk::qp14 * p0.z + ghrk::qn14 * pn.z + h * (ghrk::qa14_1 * A[1].z + ghrk::qa14_4 * A[4].z + ghrk::qa14_5 * A[5].z + ghrk::qa14_6 * A[6].z + ghrk::qa14_7 * A[7].z + ghrk::qa14_8 * A[8].z + ghrk::qa14_9 * A[9].z + ghrk::qa14_10 * A[10].z + ghrk::qa14_11 * A[11].z));
  return Q;
}

// position argument of stage 15
__device__ __forceinline__ GhVec3 gh_rkn_q15(double h, const GhVec3& q0, const GhVec3& p0,
                                             const GhVec3& pn, const GhVec3* A) {
  (void)pn;
  GhVec3 Q;
  Q.x = q0.x + h * (ghrk::qp15 * p0.x + ghrk::qn15 * pn.x + h * (ghrk::qa15_1 * A[1].x + ghrk::qa15_4 * A[4].x + ghrk::qa15_5 * A[5].x + ghrk::qa15_6 * A[6].x + ghrk::qa15_7 * A[7].x + ghrk::qa15_8 * A[8].x + ghrk::qa15_9 * A[9].x + ghrk::qa15_10 * A[10].x + ghrk::qa15_11 * A[11].x + ghrk::qa15_12 * A[12].x + ghrk::qa15_13 * A[13].x));
  Q.y = q0.y + h * (ghrk::qp15 * p0.y + ghrk::qn15 * pn.y + h * (ghrk::qa15_1 * A[1].y + ghrk::qa15_4 * A[4].y + ghrk::qa15_5 * A[5].y + ghrk::qa15_6 * A[6].y + ghrk::qa15_7 * A[7].y + ghrk::qa15_8 * A[8].y + ghrk::qa15_9 * A[9].y + ghrk::qa15_10 * A[10].y + ghrk::qa15_11 * A[11].y + ghrk::qa15_12 * A[12].y + ghrk::qa15_13 * A[13].y));
  Q.z = q0.z + h * (ghrk::qp15 * p0.z + ghrk::qn15 * pn.z + h * (ghrk::qa15_1 * A[1].z + ghrk::qa15_4 * A[4].z + ghrk::qa15_5 * A[5].z + ghrk::qa15_6 * A[6].z + ghrk::qa15_7 * A[7].z + ghrk::qa15_8 * A[8].z + ghrk::qa15_9 * A[9].z + ghrk::qa15_10 * A[10].z + ghrk::qa15_11 * A[11].z + ghrk::qa15_12 * A[12].z + ghrk::qa15_13 * A[13].z));
  return Q;
}

// position argument of stage 16
__device__ __forceinline__ GhVec3 gh_rkn_q16(double h, const GhVec3& q0, const GhVec3& p0,
                                             const GhVec3& pn, const GhVec3* A) {
  (void)pn;
  GhVec3 Q;
  Q.x = q0.x + h * (ghrk::qp16 * p0.x + ghrk::qn16 * pn.x + h * (ghrk::qa16_1 * A[1].x + ghrk::qa16_4 * A[4].x + ghrk::qa16_5 * A[5].x + ghrk::qa16_6 * A[6].x + ghrk::qa16_7 * A[7].x + ghrk::qa16_8 * A[8].x + ghrk::qa16_9 * A[9].x + ghrk::qa16_10 * A[10].x + ghrk::qa16_11 * A[11].x + ghrk::qa16_12 * A[12].x + ghrk::qa16_13 * A[13].x + ghrk::qa16_14 * A[14].x));
  Q.y = q0.y + h * (ghrk::qp16 * p0.y + ghrk::qn16 * pn.y + h * (ghrk::qa16_1 * A[1].y + ghrk::qa16_4 * A[4].y + ghrk::qa16_5 * A[5].y + ghrk::qa16_6 * A[6].y + ghrk::qa16_7 * A[7].y + ghrk::qa16_8 * A[8].y + ghrk::qa16_9 * A[9].y + ghrk::qa16_10 * A[10].y + ghrk::qa16_11 * A[11].y + ghrk::qa16_12 * A[12].y + ghrk::qa16_13 * A[13].y + ghrk::qa16_14 * A[14].y));
  Q.z = q0.z + h * (ghrk::qp16 * p0.z + ghrk::qn16 * pn.z + h * (ghrk::qa16_1 * A[1].z + ghrk::qa16_4 * A[4].z + ghrk::qa16_5 * A[5].z + ghrk::qa16_6 * A[6].z + ghrk::qa16_7 * A[7].z + ghrk::qa16_8 * A[8].z + ghrk::qa16_9 * A[9].z + ghrk::qa16_10 * A[10].z + ghrk::qa16_11 * A[11].z + ghrk::qa16_12 * A[12].z + ghrk::qa16_13 * A[13].z + ghrk::qa16_14 * A[14].z));
  return Q;
}

// New state (q_new, p_new) and the two error sums of leung_dop853.c:376-404:
// per component, sk = atol + rtol max(|y0|, |y_new|), err3 += (e3/sk)^2,
// err5 += (e5/sk)^2, with e3 = sum (b - bhh) k, e5 = sum er k (one
// reciprocal of sk per component; the norm only steers the step size).
__device__ __forceinline__ void gh_rkn_solution(double h, const GhVec3& q0, const GhVec3& p0,
                                               const GhVec3* A, double atol, double rtol,
                                               GhVec3& qn, GhVec3& pn, double& err5,
                                               double& err3) {
  err5 = 0.0;
  err3 = 0.0;
  {
    const double kq = ghrk::bp * p0.x + h * (ghrk::ba_1 * A[1].x + ghrk::ba_4 * A[4].x + ghrk::ba_5 * A[5].x + ghrk::ba_6 * A[6].x + ghrk::ba_7 * A[7].x + ghrk::ba_8 * A[8].x + ghrk::ba_9 * A[9].x + ghrk::ba_10 * A[10].x + ghrk::ba_11 * A[11].x);
    qn.x = q0.x + h * kq;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(q0.x), fabs(qn.x)));
    double sqr = (ghrk::e3p * p0.x + h * (ghrk::e3a_1 * A[1].x + ghrk::e3a_4 * A[4].x + ghrk::e3a_5 * A[5].x + ghrk::e3a_6 * A[6].x + ghrk::e3a_7 * A[7].x + ghrk::e3a_8 * A[8].x + ghrk::e3a_9 * A[9].x + ghrk::e3a_10 * A[10].x + ghrk::e3a_11 * A[11].x)) * isk;
    err3 += sqr * sqr;
    sqr = (ghrk::e5p * p0.x + h * (ghrk::e5a_1 * A[1].x + ghrk::e5a_4 * A[4].x + ghrk::e5a_5 * A[5].x + ghrk::e5a_6 * A[6].x + ghrk::e5a_7 * A[7].x + ghrk::e5a_8 * A[8].x + ghrk::e5a_9 * A[9].x + ghrk::e5a_10 * A[10].x + ghrk::e5a_11 * A[11].x)) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kq = ghrk::bp * p0.y + h * (ghrk::ba_1 * A[1].y + ghrk::ba_4 * A[4].y + ghrk::ba_5 * A[5].y + ghrk::ba_6 * A[6].y + ghrk::ba_7 * A[7].y + ghrk::ba_8 * A[8].y + ghrk::ba_9 * A[9].y + ghrk::ba_10 * A[10].y + ghrk::ba_11 * A[11].y);
    qn.y = q0.y + h * kq;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(q0.y), fabs(qn.y)));
    double sqr = (ghrk::e3p * p0.y + h * (ghrk::e3a_1 * A[1].y + ghrk::e3a_4 * A[4].y + ghrk::e3a_5 * A[5].y + ghrk::e3a_6 * A[6].y + ghrk::e3a_7 * A[7].y + ghrk::e3a_8 * A[8].y + ghrk::e3a_9 * A[9].y + ghrk::e3a_10 * A[10].y + ghrk::e3a_11 * A[11].y)) * isk;
    err3 += sqr * sqr;
    sqr = (ghrk::e5p * p0.y + h * (ghrk::e5a_1 * A[1].y + ghrk::e5a_4 * A[4].y + ghrk::e5a_5 * A[5].y + ghrk::e5a_6 * A[6].y + ghrk::e5a_7 * A[7].y + ghrk::e5a_8 * A[8].y + ghrk::e5a_9 * A[9].y + ghrk::e5a_10 * A[10].y + ghrk::e5a_11 * A[11].y)) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kq = ghrk::bp * p0.z + h * (ghrk::ba_1 * A[1].z + ghrk::ba_4 * A[4].z + ghrk::ba_5 * A[5].z + ghrk::ba_6 * A[6].z + ghrk::ba_7 * A[7].z + ghrk::ba_8 * A[8].z + ghrk::ba_9 * A[9].z + ghrk::ba_10 * A[10].z + ghrk::ba_11 * A[11].z);
    qn.z = q0.z + h * kq;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(q0.z), fabs(qn.z)));
    double sqr = (ghrk::e3p * p0.z + h * (ghrk::e3a_1 * A[1].z + ghrk::e3a_4 * A[4].z + ghrk::e3a_5 * A[5].z + ghrk::e3a_6 * A[6].z + ghrk::e3a_7 * A[7].z + ghrk::e3a_8 * A[8].z + ghrk::e3a_9 * A[9].z + ghrk::e3a_10 * A[10].z + ghrk::e3a_11 * A[11].z)) * isk;
    err3 += sqr * sqr;
    sqr = (ghrk::e5p * p0.z + h * (ghrk::e5a_1 * A[1].z + ghrk::e5a_4 * A[4].z + ghrk::e5a_5 * A[5].z + ghrk::e5a_6 * A[6].z + ghrk::e5a_7 * A[7].z + ghrk::e5a_8 * A[8].z + ghrk::e5a_9 * A[9].z + ghrk::e5a_10 * A[10].z + ghrk::e5a_11 * A[11].z)) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kp = ghdp::b1 * A[1].x + ghdp::b6 * A[6].x + ghdp::b7 * A[7].x + ghdp::b8 * A[8].x + ghdp::b9 * A[9].x + ghdp::b10 * A[10].x + ghdp::b11 * A[11].x + ghdp::b12 * A[12].x;
    pn.x = p0.x + h * kp;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(p0.x), fabs(pn.x)));
    double sqr = (kp - ghdp::bhh1 * A[1].x - ghdp::bhh2 * A[9].x - ghdp::bhh3 * A[12].x) * isk;
    err3 += sqr * sqr;
    sqr = (ghdp::er1 * A[1].x + ghdp::er6 * A[6].x + ghdp::er7 * A[7].x + ghdp::er8 * A[8].x + ghdp::er9 * A[9].x + ghdp::er10 * A[10].x + ghdp::er11 * A[11].x + ghdp::er12 * A[12].x) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kp = ghdp::b1 * A[1].y + ghdp::b6 * A[6].y + ghdp::b7 * A[7].y + ghdp::b8 * A[8].y + ghdp::b9 * A[9].y + ghdp::b10 * A[10].y + ghdp::b11 * A[11].y + ghdp::b12 * A[12].y;
    pn.y = p0.y + h * kp;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(p0.y), fabs(pn.y)));
    double sqr = (kp - ghdp::bhh1 * A[1].y - ghdp::bhh2 * A[9].y - ghdp::bhh3 * A[12].y) * isk;
    err3 += sqr * sqr;
    sqr = (ghdp::er1 * A[1].y + ghdp::er6 * A[6].y + ghdp::er7 * A[7].y + ghdp::er8 * A[8].y + ghdp::er9 * A[9].y + ghdp::er10 * A[10].y + ghdp::er11 * A[11].y + ghdp::er12 * A[12].y) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kp = ghdp::b1 * A[1].z + ghdp::b6 * A[6].z + ghdp::b7 * A[7].z + ghdp::b8 * A[8].z + ghdp::b9 * A[9].z + ghdp::b10 * A[10].z + ghdp::b11 * A[11].z + ghdp::b12 * A[12].z;
    pn.z = p0.z + h * kp;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(p0.z), fabs(pn.z)));
    double sqr = (kp - ghdp::bhh1 * A[1].z - ghdp::bhh2 * A[9].z - ghdp::bhh3 * A[12].z) * isk;
    err3 += sqr * sqr;
    sqr = (ghdp::er1 * A[1].z + ghdp::er6 * A[6].z + ghdp::er7 * A[7].z + ghdp::er8 * A[8].z + ghdp::er9 * A[9].z + ghdp::er10 * A[10].z + ghdp::er11 * A[11].z + ghdp::er12 * A[12].z) * isk;
    err5 += sqr * sqr;
  }
}

// Dense output (leung_dop853.c:430-474) of the accepted step at the time
// t_old + sd h (s1 = 1 - sd), after the dense stages 14-16: per component
// the coefficients rc1..rc8 of the continuous extension, then
// rc1 + sd (rc2 + s1 (rc3 + sd (rc4 + s1 (rc5 + sd (rc6 + s1 (rc7 + sd rc8)))))).
__device__ __forceinline__ void gh_rkn_dense_eval(double h, double sd, double s1, const GhVec3& q0,
                                                 const GhVec3& p0, const GhVec3& qn,
                                                 const GhVec3& pn, const GhVec3* A, double* yo) {
  {
    const double ydiff = qn.x - q0.x;
    const double bspl = h * p0.x - ydiff;
    const double r0 = q0.x, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * pn.x - bspl;
    const double r4 = h * (ghrk::dp4 * p0.x + ghrk::dn4 * pn.x + h * (ghrk::da4_1 * A[1].x + ghrk::da4_4 * A[4].x + ghrk::da4_5 * A[5].x + ghrk::da4_6 * A[6].x + ghrk::da4_7 * A[7].x + ghrk::da4_8 * A[8].x + ghrk::da4_9 * A[9].x + ghrk::da4_10 * A[10].x + ghrk::da4_11 * A[11].x + ghrk::da4_12 * A[12].x + ghrk::da4_13 * A[13].x + ghrk::da4_14 * A[14].x + ghrk::da4_15 * A[15].x));
    const double r5 = h * (ghrk::dp5 * p0.x + ghrk::dn5 * pn.x + h * (ghrk::da5_1 * A[1].x + ghrk::da5_4 * A[4].x + ghrk::da5_5 * A[5].x + ghrk::da5_6 * A[6].x + ghrk::da5_7 * A[7].x + ghrk::da5_8 * A[8].x + ghrk::da5_9 * A[9].x + ghrk::da5_10 * A[10].x + ghrk::da5_11 * A[11].x + ghrk::da5_12 * A[12].x + ghrk::da5_13 * A[13].x + ghrk::da5_14 * A[14].x + ghrk::da5_15 * A[15].x));
    const double r6 = h * (ghrk::dp6 * p0.x + ghrk::dn6 * pn.x + h * (ghrk::da6_1 * A[1].x + ghrk::da6_4 * A[4].x + ghrk::da6_5 * A[5].x + ghrk::da6_6 * A[6].x + ghrk::da6_7 * A[7].x + ghrk::da6_8 * A[8].x + ghrk::da6_9 * A[9].x + ghrk::da6_10 * A[10].x + ghrk::da6_11 * A[11].x + ghrk::da6_12 * A[12].x + ghrk::da6_13 * A[13].x + ghrk::da6_14 * A[14].x + ghrk::da6_15 * A[15].x));
    const double r7 = h * (ghrk::dp7 * p0.x + ghrk::dn7 * pn.x + h * (ghrk::da7_1 * A[1].x + ghrk::da7_4 * A[4].x + ghrk::da7_5 * A[5].x + ghrk::da7_6 * A[6].x + ghrk::da7_7 * A[7].x + ghrk::da7_8 * A[8].x + ghrk::da7_9 * A[9].x + ghrk::da7_10 * A[10].x + ghrk::da7_11 * A[11].x + ghrk::da7_12 * A[12].x + ghrk::da7_13 * A[13].x + ghrk::da7_14 * A[14].x + ghrk::da7_15 * A[15].x));
    yo[0] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = qn.y - q0.y;
    const double bspl = h * p0.y - ydiff;
    const double r0 = q0.y, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * pn.y - bspl;
    const double r4 = h * (ghrk::dp4 * p0.y + ghrk::dn4 * pn.y + h * (ghrk::da4_1 * A[1].y + ghrk::da4_4 * A[4].y + ghrk::da4_5 * A[5].y + ghrk::da4_6 * A[6].y + ghrk::da4_7 * A[7].y + ghrk::da4_8 * A[8].y + ghrk::da4_9 * A[9].y + ghrk::da4_10 * A[10].y + ghrk::da4_11 * A[11].y + ghrk::da4_12 * A[12].y + ghrk::da4_13 * A[13].y + ghrk::da4_14 * A[14].y + ghrk::da4_15 * A[15].y));
    const double r5 = h * (ghrk::dp5 * p0.y + ghrk::dn5 * pn.y + h * (ghrk::da5_1 * A[1].y + ghrk::da5_4 * A[4].y + ghrk::da5_5 * A[5].y + ghrk::da5_6 * A[6].y + ghrk::da5_7 * A[7].y + ghrk::da5_8 * A[8].y + ghrk::da5_9 * A[9].y + ghrk::da5_10 * A[10].y + ghrk::da5_11 * A[11].y + ghrk::da5_12 * A[12].y + ghrk::da5_13 * A[13].y + ghrk::da5_14 * A[14].y + ghrk::da5_15 * A[15].y));
    const double r6 = h * (ghrk::dp6 * p0.y + ghrk::dn6 * pn.y + h * (ghrk::da6_1 * A[1].y + ghrk::da6_4 * A[4].y + ghrk::da6_5 * A[5].y + ghrk::da6_6 * A[6].y + ghrk::da6_7 * A[7].y + ghrk::da6_8 * A[8].y + ghrk::da6_9 * A[9].y + ghrk::da6_10 * A[10].y + ghrk::da6_11 * A[11].y + ghrk::da6_12 * A[12].y + ghrk::da6_13 * A[13].y + ghrk::da6_14 * A[14].y + ghrk::da6_15 * A[15].y));
    const double r7 = h * (ghrk::dp7 * p0.y + ghrk::dn7 * pn.y + h * (ghrk::da7_1 * A[1].y + ghrk::da7_4 * A[4].y + ghrk::da7_5 * A[5].y + ghrk::da7_6 * A[6].y + ghrk::da7_7 * A[7].y + ghrk::da7_8 * A[8].y + ghrk::da7_9 * A[9].y + ghrk::da7_10 * A[10].y + ghrk::da7_11 * A[11].y + ghrk::da7_12 * A[12].y + ghrk::da7_13 * A[13].y + ghrk::da7_14 * A[14].y + ghrk::da7_15 * A[15].y));
    yo[1] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = qn.z - q0.z;
    const double bspl = h * p0.z - ydiff;
    const double r0 = q0.z, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * pn.z - bspl;
    const double r4 = h * (ghrk::dp4 * p0.z + ghrk::dn4 * pn.z + h * (ghrk::da4_1 * A[1].z + ghrk::da4_4 * A[4].z + ghrk::da4_5 * A[5].z + ghrk::da4_6 * A[6].z + ghrk::da4_7 * A[7].z + ghrk::da4_8 * A[8].z + ghrk::da4_9 * A[9].z + ghrk::da4_10 * A[10].z + ghrk::da4_11 * A[11].z + ghrk::da4_12 * A[12].z + ghrk::da4_13 * A[13].z + ghrk::da4_14 * A[14].z + ghrk::da4_15 * A[15].z));
    const double r5 = h * (ghrk::dp5 * p0.z + ghrk::dn5 * pn.z + h * (ghrk::da5_1 * A[1].z + ghrk::da5_4 * A[4].z + ghrk::da5_5 * A[5].z + ghrk::da5_6 * A[6].z + ghrk::da5_7 * A[7].z + ghrk::da5_8 * A[8].z + ghrk::da5_9 * A[9].z + ghrk::da5_10 * A[10].z + ghrk::da5_11 * A[11].z + ghrk::da5_12 * A[12].z + ghrk::da5_13 * A[13].z + ghrk::da5_14 * A[14].z + ghrk::da5_15 * A[15].z));
    const double r6 = h * (ghrk::dp6 * p0.z + ghrk::dn6 * pn.z + h * (ghrk::da6_1 * A[1].z + ghrk::da6_4 * A[4].z + ghrk::da6_5 * A[5].z + ghrk::da6_6 * A[6].z + ghrk::da6_7 * A[7].z + ghrk::da6_8 * A[8].z + ghrk::da6_9 * A[9].z + ghrk::da6_10 * A[10].z + ghrk::da6_11 * A[11].z + ghrk::da6_12 * A[12].z + ghrk::da6_13 * A[13].z + ghrk::da6_14 * A[14].z + ghrk::da6_15 * A[15].z));
    const double r7 = h * (ghrk::dp7 * p0.z + ghrk::dn7 * pn.z + h * (ghrk::da7_1 * A[1].z + ghrk::da7_4 * A[4].z + ghrk::da7_5 * A[5].z + ghrk::da7_6 * A[6].z + ghrk::da7_7 * A[7].z + ghrk::da7_8 * A[8].z + ghrk::da7_9 * A[9].z + ghrk::da7_10 * A[10].z + ghrk::da7_11 * A[11].z + ghrk::da7_12 * A[12].z + ghrk::da7_13 * A[13].z + ghrk::da7_14 * A[14].z + ghrk::da7_15 * A[15].z));
    yo[2] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = pn.x - p0.x;
    const double bspl = h * A[1].x - ydiff;
    const double r0 = p0.x, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * A[13].x - bspl;
    const double r4 = h * (ghdp::d41 * A[1].x + ghdp::d46 * A[6].x + ghdp::d47 * A[7].x + ghdp::d48 * A[8].x + ghdp::d49 * A[9].x + ghdp::d410 * A[10].x + ghdp::d411 * A[11].x + ghdp::d412 * A[12].x + ghdp::d413 * A[13].x + ghdp::d414 * A[14].x + ghdp::d415 * A[15].x + ghdp::d416 * A[16].x);
    const double r5 = h * (ghdp::d51 * A[1].x + ghdp::d56 * A[6].x + ghdp::d57 * A[7].x + ghdp::d58 * A[8].x + ghdp::d59 * A[9].x + ghdp::d510 * A[10].x + ghdp::d511 * A[11].x + ghdp::d512 * A[12].x + ghdp::d513 * A[13].x + ghdp::d514 * A[14].x + ghdp::d515 * A[15].x + ghdp::d516 * A[16].x);
    const double r6 = h * (ghdp::d61 * A[1].x + ghdp::d66 * A[6].x + ghdp::d67 * A[7].x + ghdp::d68 * A[8].x + ghdp::d69 * A[9].x + ghdp::d610 * A[10].x + ghdp::d611 * A[11].x + ghdp::d612 * A[12].x + ghdp::d613 * A[13].x + ghdp::d614 * A[14].x + ghdp::d615 * A[15].x + ghdp::d616 * A[16].x);
    const double r7 = h * (ghdp::d71 * A[1].x + ghdp::d76 * A[6].x + ghdp::d77 * A[7].x + ghdp::d78 * A[8].x + ghdp::d79 * A[9].x + ghdp::d710 * A[10].x + ghdp::d711 * A[11].x + ghdp::d712 * A[12].x + ghdp::d713 * A[13].x + ghdp::d714 * A[14].x + ghdp::d715 * A[15].x + ghdp::d716 * A[16].x);
    yo[3] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = pn.y - p0.y;
    const double bspl = h * A[1].y - ydiff;
    const double r0 = p0.y, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * A[13].y - bspl;
    const double r4 = h * (ghdp::d41 * A[1].y + ghdp::d46 * A[6].y + ghdp::d47 * A[7].y + ghdp::d48 * A[8].y + ghdp::d49 * A[9].y + ghdp::d410 * A[10].y + ghdp::d411 * A[11].y + ghdp::d412 * A[12].y + ghdp::d413 * A[13].y + ghdp::d414 * A[14].y + ghdp::d415 * A[15].y + ghdp::d416 * A[16].y);
    const double r5 = h * (ghdp::d51 * A[1].y + ghdp::d56 * A[6].y + ghdp::d57 * A[7].y + ghdp::d58 * A[8].y + ghdp::d59 * A[9].y + ghdp::d510 * A[10].y + ghdp::d511 * A[11].y + ghdp::d512 * A[12].y + ghdp::d513 * A[13].y + ghdp::d514 * A[14].y + ghdp::d515 * A[15].y + ghdp::d516 * A[16].y);
    const double r6 = h * (ghdp::d61 * A[1].y + ghdp::d66 * A[6].y + ghdp::d67 * A[7].y + ghdp::d68 * A[8].y + ghdp::d69 * A[9].y + ghdp::d610 * A[10].y + ghdp::d611 * A[11].y + ghdp::d612 * A[12].y + ghdp::d613 * A[13].y + ghdp::d614 * A[14].y + ghdp::d615 * A[15].y + ghdp::d616 * A[16].y);
    const double r7 = h * (ghdp::d71 * A[1].y + ghdp::d76 * A[6].y + ghdp::d77 * A[7].y + ghdp::d78 * A[8].y + ghdp::d79 * A[9].y + ghdp::d710 * A[10].y + ghdp::d711 * A[11].y + ghdp::d712 * A[12].y + ghdp::d713 * A[13].y + ghdp::d714 * A[14].y + ghdp::d715 * A[15].y + ghdp::d716 * A[16].y);
    yo[4] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = pn.z - p0.z;
    const double bspl = h * A[1].z - ydiff;
    const double r0 = p0.z, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * A[13].z - bspl;
    const double r4 = h * (ghdp::d41 * A[1].z + ghdp::d46 * A[6].z + ghdp::d47 * A[7].z + ghdp::d48 * A[8].z + ghdp::d49 * A[9].z + ghdp::d410 * A[10].z + ghdp::d411 * A[11].z + ghdp::d412 * A[12].z + ghdp::d413 * A[13].z + ghdp::d414 * A[14].z + ghdp::d415 * A[15].z + ghdp::d416 * A[16].z);
    const double r5 = h * (ghdp::d51 * A[1].z + ghdp::d56 * A[6].z + ghdp::d57 * A[7].z + ghdp::d58 * A[8].z + ghdp::d59 * A[9].z + ghdp::d510 * A[10].z + ghdp::d511 * A[11].z + ghdp::d512 * A[12].z + ghdp::d513 * A[13].z + ghdp::d514 * A[14].z + ghdp::d515 * A[15].z + ghdp::d516 * A[16].z);
    const double r6 = h * (ghdp::d61 * A[1].z + ghdp::d66 * A[6].z + ghdp::d67 * A[7].z + ghdp::d68 * A[8].z + ghdp::d69 * A[9].z + ghdp::d610 * A[10].z + ghdp::d611 * A[11].z + ghdp::d612 * A[12].z + ghdp::d613 * A[13].z + ghdp::d614 * A[14].z + ghdp::d615 * A[15].z + ghdp::d616 * A[16].z);
    const double r7 = h * (ghdp::d71 * A[1].z + ghdp::d76 * A[6].z + ghdp::d77 * A[7].z + ghdp::d78 * A[8].z + ghdp::d79 * A[9].z + ghdp::d710 * A[10].z + ghdp::d711 * A[11].z + ghdp::d712 * A[12].z + ghdp::d713 * A[13].z + ghdp::d714 * A[14].z + ghdp::d715 * A[15].z + ghdp::d716 * A[16].z);
    yo[5] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
}
// position argument of stage 2
__device__ __forceinline__ GhVec2 gh_rkn_q2(double h, const GhVec2& q0, const GhVec2& p0,
                                             const GhVec2& pn, const GhVec2* A) {
  (void)pn;
  GhVec2 Q;
  Q.x = q0.x + h * (ghrk::qp2 * p0.x);
  Q.y = q0.y + h * (ghrk::qp2 * p0.y);
  return Q;
}

// position argument of stage 3
__device__ __forceinline__ GhVec2 gh_rkn_q3(double h, const GhVec2& q0, const GhVec2& p0,
                                             const GhVec2& pn, const GhVec2* A) {
  (void)pn;
  GhVec2 Q;
  Q.x = q0.x + h * (ghrk::qp3 * p0.x + h * (ghrk::qa3_1 * A[1].x));
  Q.y = q0.y + h * (ghrk::qp3 * p0.y + h * (ghrk::qa3_1 * A[1].y));
  return Q;
}

// position argument of stage 4
__device__ __forceinline__ GhVec2 gh_rkn_q4(double h, const GhVec2& q0, const GhVec2& p0,
                                             const GhVec2& pn, const GhVec2* A) {
  (void)pn;
  GhVec2 Q;
  Q.x = q0.x + h * (ghrk::qp4 * p0.x + h * (ghrk::qa4_1 * A[1].x + ghrk::qa4_2 * A[2].x));
  Q.y = q0.y + h * (ghrk::qp4 * p0.y + h * (ghrk::qa4_1 * A[1].y + ghrk::qa4_2 * A[2].y));
  return Q;
}

// position argument of stage 5
__device__ __forceinline__ GhVec2 gh_rkn_q5(double h, const GhVec2& q0, const GhVec2& p0,
                                             const GhVec2& pn, const GhVec2* A) {
  (void)pn;
  GhVec2 Q;
  Q.x = q0.x + h * (ghrk::qp5 * p0.x + h * (ghrk::qa5_1 * A[1].x + ghrk::qa5_2 * A[2].x + ghrk::qa5_3 * A[3].x));
  Q.y = q0.y + h * (ghrk::qp5 * p0.y + h * (ghrk::qa5_1 * A[1].y + ghrk::qa5_2 * A[2].y + ghrk::qa5_3 * A[3].y));
  return Q;
}

// position argument of stage 6
__device__ __forceinline__ GhVec2 gh_rkn_q6(double h, const GhVec2& q0, const GhVec2& p0,
                                             const GhVec2& pn, const GhVec2* A) {
  (void)pn;
  GhVec2 Q;
  Q.x = q0.x + h * (ghrk::qp6 * p0.x + h * (ghrk::qa6_1 * A[1].x + ghrk::qa6_3 * A[3].x + ghrk::qa6_4 * A[4].x));
  Q.y = q0.y + h * (ghrk::qp6 * p0.y + h * (ghrk::qa6_1 * A[1].y + ghrk::qa6_3 * A[3].y + ghrk::qa6_4 * A[4].y));
  return Q;
}

// position argument of stage 7
__device__ __forceinline__ GhVec2 gh_rkn_q7(double h, const GhVec2& q0, const GhVec2& p0,
                                             const GhVec2& pn, const GhVec2* A) {
  (void)pn;
  GhVec2 Q;
  Q.x = q0.x + h * (ghrk::qp7 * p0.x + h * (ghrk::qa7_1 * A[1].x + ghrk::qa7_3 * A[3].x + ghrk::qa7_4 * A[4].x + ghrk::qa7_5 * A[5].x));
  Q.y = q0.y + h * (ghrk::qp7 * p0.y + h * (ghrk::qa7_1 * A[1].y + ghrk::qa7_3 * A[3].y + ghrk::qa7_4 * A[4].y + ghrk::qa7_5 * A[5].y));
  return Q;
}

// position argument of stage 8
__device__ __forceinline__ GhVec2 gh_rkn_q8(double h, const GhVec2& q0, const GhVec2& p0,
                                             const GhVec2& pn, const GhVec2* A) {
  (void)pn;
  GhVec2 Q;
  Q.x = q0.x + h * (ghrk::qp8 * p0.x + h * (ghrk::qa8_1 * A[1].x + ghrk::qa8_3 * A[3].x + ghrk::qa8_4 * A[4].x + ghrk::qa8_5 * A[5].x + ghrk::qa8_6 * A[6].x));
  Q.y = q0.y + h * (ghrk::qp8 * p0.y + h * (ghrk::qa8_1 * A[1].y + ghrk::qa8_3 * A[3].y + ghrk::qa8_4 * A[4].y + ghrk::qa8_5 * A[5].y + ghrk::qa8_6 * A[6].y));
  return Q;
}

// position argument of stage 9
__device__ __forceinline__ GhVec2 gh_rkn_q9(double h, const GhVec2& q0, const GhVec2& p0,
                                             const GhVec2& pn, const GhVec2* A) {
  (void)pn;
  GhVec2 Q;
  Q.x = q0.x + h * (ghrk::qp9 * p0.x + h * (ghrk::qa9_1 * A[1].x + ghrk::qa9_3 * A[3].x + ghrk::qa9_4 * A[4].x + ghrk::qa9_5 * A[5].x + ghrk::qa9_6 * A[6].x + ghrk::qa9_7 * A[7].x));
  Q.y = q0.y + h * (ghrk::qp9 * p0.y + h * (ghrk::qa9_1 * A[1].y + ghrk::qa9_3 * A[3].y + ghrk::qa9_4 * A[4].y + ghrk::qa9_5 * A[5].y + ghrk::qa9_6 * A[6].y + ghrk::qa9_7 * A[7].y));
  return Q;
}

// position argument of stage 10
__device__ __forceinline__ GhVec2 gh_rkn_q10(double h, const GhVec2& q0, const GhVec2& p0,
                                             const GhVec2& pn, const GhVec2* A) {
  (void)pn;
  GhVec2 Q;
  Q.x = q0.x + h * (ghrk::qp10 * p0.x + h * (ghrk::qa10_1 * A[1].x + ghrk::qa10_3 * A[3].x + ghrk::qa10_4 * A[4].x + ghrk::qa10_5 * A[5].x + ghrk::qa10_6 * A[6].x + ghrk::qa10_7 * A[7].x + ghrk::qa10_8 * A[8].x));
  Q.y = q0.y + h * (ghrk::qp10 * p0.y + h * (ghrk::qa10_1 * A[1].y + ghrk::qa10_3 * A[3].y + ghrk::qa10_4 * A[4].y + ghrk::qa10_5 * A[5].y + ghrk::qa10_6 * A[6].y + ghrk::qa10_7 * A[7].y + ghrk::qa10_8 * A[8].y));
  return Q;
}

// position argument of stage 11
__device__ __forceinline__ GhVec2 gh_rkn_q11(double h, const GhVec2& q0, const GhVec2& p0,
                                             const GhVec2& pn, const GhVec2* A) {
  (void)pn;
  GhVec2 Q;
  Q.x = q0.x + h * (ghrk::qp11 * p0.x + h * (ghrk::qa11_1 * A[1].x + ghrk::qa11_3 * A[3].x + ghrk::qa11_4 * A[4].x + ghrk::qa11_5 * A[5].x + ghrk::qa11_6 * A[6].x + ghrk::qa11_7 * A[7].x + ghrk::qa11_8 * A[8].x + ghrk::qa11_9 * A[9].x));
  Q.y = q0.y + h * (ghrk::qp11 * p0.y + h * (ghrk::qa11_1 * A[1].y + ghrk::qa11_3 * A[3].y + ghrk::qa11_4 * A[4].y + ghrk::qa11_5 * A[5].y + ghrk::qa11_6 * A[6].y + ghrk::qa11_7 * A[7].y + ghrk::qa11_8 * A[8].y + ghrk::qa11_9 * A[9].y));
  return Q;
}

// position argument of stage 12
__device__ __forceinline__ GhVec2 gh_rkn_q12(double h, const GhVec2& q0, const GhVec2& p0,
                                             const GhVec2& pn, const GhVec2* A) {
  (void)pn;
  GhVec2 Q;
  Q.x = q0.x + h * (ghrk::qp12 * p0.x + h * (ghrk::qa12_1 * A[1].x + ghrk::qa12_3 * A[3].x + ghrk::qa12_4 * A[4].x + ghrk::qa12_5 * A[5].x + ghrk::qa12_6 * A[6].x + ghrk::qa12_7 * A[7].x + ghrk::qa12_8 * A[8].x + ghrk::qa12_9 * A[9].x + ghrk::qa12_10 * A[10].x));
  Q.y = q0.y + h * (ghrk::qp12 * p0.y + h * (ghrk::qa12_1 * A[1].y + ghrk::qa12_3 * A[3].y + ghrk::qa12_4 * A[4].y + ghrk::qa12_5 * A[5].y + ghrk::qa12_6 * A[6].y + ghrk::qa12_7 * A[7].y + ghrk::qa12_8 * A[8].y + ghrk::qa12_9 * A[9].y + ghrk::qa12_10 * A[10].y));
  return Q;
}

// position argument of stage 14
__device__ __forceinline__ GhVec2 gh_rkn_q14(double h, const GhVec2& q0, const GhVec2& p0,
                                             const GhVec2& pn, const GhVec2* A) {
  (void)pn;
  GhVec2 Q;
  Q.x = q0.x + h * (ghrk::qp14 * p0.x + ghrk::qn14 * pn.x + h * (ghrk::qa14_1 * A[1].x + ghrk::qa14_4 * A[4].x + ghrk::qa14_5 * A[5].x + ghrk::qa14_6 * A[6].x + ghrk::qa14_7 * A[7].x + ghrk::qa14_8 * A[8].x + ghrk::qa14_9 * A[9].x + ghrk::qa14_10 * A[10].x + ghrk::qa14_11 * A[11].x));
  Q.y = q0.y + h * (ghrk::qp14 * p0.y + ghrk::qn14 * pn.y + h * (ghrk::qa14_1 * A[1].y + ghrk::qa14_4 * A[4].y + ghrk::qa14_5 * A[5].y + ghrk::qa14_6 * A[6].y + ghrk::qa14_7 * A[7].y + ghrk::qa14_8 * A[8].y + ghrk::qa14_9 * A[9].y + ghrk::qa14_10 * A[10].y + ghrk::qa14_11 * A[11].y));
  return Q;
}

// position argument of stage 15
__device__ __forceinline__ GhVec2 gh_rkn_q15(double h, const GhVec2& q0, const GhVec2& p0,
                                             const GhVec2& pn, const GhVec2* A) {
  (void)pn;
  GhVec2 Q;
  Q.x = q0.x + h * (ghrk::qp15 * p0.x + ghrk::qn15 * pn.x + h * (ghrk::qa15_1 * A[1].x + ghrk::qa15_4 * A[4].x + ghrk::qa15_5 * A[5].x + ghrk::qa15_6 * A[6].x + ghrk::qa15_7 * A[7].x + ghrk::qa15_8 * A[8].x + ghrk::qa15_9 * A[9].x + ghrk::qa15_10 * A[10].x + ghrk::qa15_11 * A[11].x + ghrk::qa15_12 * A[12].x + ghrk::qa15_13 * A[13].x));
  Q.y = q0.y + h * (ghrk::qp15 * p0.y + ghrk::qn15 * pn.y + h * (ghrk::qa15_1 * A[1].y + ghrk::qa15_4 * A[4].y + ghrk::qa15_5 * A[5].y + ghrk::qa15_6 * A[6].y + ghrk::qa15_7 * A[7].y + ghrk::qa15_8 * A[8].y + ghrk::qa15_9 * A[9].y + ghrk::qa15_10 * A[10].y + ghrk::qa15_11 * A[11].y + ghrk::qa15_12 * A[12].y + ghrk::qa15_13 * A[13].y));
  return Q;
}

// position argument of stage 16
__device__ __forceinline__ GhVec2 gh_rkn_q16(double h, const GhVec2& q0, const GhVec2& p0,
                                             const GhVec2& pn, const GhVec2* A) {
  (void)pn;
  GhVec2 Q;
  Q.x = q0.x + h * (ghrk::qp16 * p0.x + ghrk::qn16 * pn.x + h * (ghrk::qa16_1 * A[1].x + ghrk::qa16_4 * A[4].x + ghrk::qa16_5 * A[5].x + ghrk::qa16_6 * A[6].x + ghrk::qa16_7 * A[7].x + ghrk::qa16_8 * A[8].x + ghrk::qa16_9 * A[9].x + ghrk::qa16_10 * A[10].x + ghrk::qa16_11 * A[11].x + ghrk::qa16_12 * A[12].x + ghrk::qa16_13 * A[13].x + ghrk::qa16_14 * A[14].x));
  Q.y = q0.y + h * (ghrk::qp16 * p0.y + ghrk::qn16 * pn.y + h * (ghrk::qa16_1 * A[1].y + ghrk::qa16_4 * A[4].y + ghrk::qa16_5 * A[5].y + ghrk::qa16_6 * A[6].y + ghrk::qa16_7 * A[7].y + ghrk::qa16_8 * A[8].y + ghrk::qa16_9 * A[9].y + ghrk::qa16_10 * A[10].y + ghrk::qa16_11 * A[11].y + ghrk::qa16_12 * A[12].y + ghrk::qa16_13 * A[13].y + ghrk::qa16_14 * A[14].y));
  return Q;
}

// New state (q_new, p_new) and the two error sums of leung_dop853.c:376-404:
// per component, sk = atol + rtol max(|y0|, |y_new|), err3 += (e3/sk)^2,
// err5 += (e5/sk)^2, with e3 = sum (b - bhh) k, e5 = sum er k (one
// reciprocal of sk per component; the norm only steers the step size).
__device__ __forceinline__ void gh_rkn_solution(double h, const GhVec2& q0, const GhVec2& p0,
                                               const GhVec2* A, double atol, double rtol,
                                               GhVec2& qn, GhVec2& pn, double& err5,
                                               double& err3) {
  err5 = 0.0;
  err3 = 0.0;
  {
    const double kq = ghrk::bp * p0.x + h * (ghrk::ba_1 * A[1].x + ghrk::ba_4 * A[4].x + ghrk::ba_5 * A[5].x + ghrk::ba_6 * A[6].x + ghrk::ba_7 * A[7].x + ghrk::ba_8 * A[8].x + ghrk::ba_9 * A[9].x + ghrk::ba_10 * A[10].x + ghrk::ba_11 * A[11].x);
    qn.x = q0.x + h * kq;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(q0.x), fabs(qn.x)));
    double sqr = (ghrk::e3p * p0.x + h * (ghrk::e3a_1 * A[1].x + ghrk::e3a_4 * A[4].x + ghrk::e3a_5 * A[5].x + ghrk::e3a_6 * A[6].x + ghrk::e3a_7 * A[7].x + ghrk::e3a_8 * A[8].x + ghrk::e3a_9 * A[9].x + ghrk::e3a_10 * A[10].x + ghrk::e3a_11 * A[11].x)) * isk;
    err3 += sqr * sqr;
    sqr = (ghrk::e5p * p0.x + h * (ghrk::e5a_1 * A[1].x + ghrk::e5a_4 * A[4].x + ghrk::e5a_5 * A[5].x + ghrk::e5a_6 * A[6].x + ghrk::e5a_7 * A[7].x + ghrk::e5a_8 * A[8].x + ghrk::e5a_9 * A[9].x + ghrk::e5a_10 * A[10].x + ghrk::e5a_11 * A[11].x)) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kq = ghrk::bp * p0.y + h * (ghrk::ba_1 * A[1].y + ghrk::ba_4 * A[4].y + ghrk::ba_5 * A[5].y + ghrk::ba_6 * A[6].y + ghrk::ba_7 * A[7].y + ghrk::ba_8 * A[8].y + ghrk::ba_9 * A[9].y + ghrk::ba_10 * A[10].y + ghrk::ba_11 * A[11].y);
    qn.y = q0.y + h * kq;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(q0.y), fabs(qn.y)));
    double sqr = (ghrk::e3p * p0.y + h * (ghrk::e3a_1 * A[1].y + ghrk::e3a_4 * A[4].y + ghrk::e3a_5 * A[5].y + ghrk::e3a_6 * A[6].y + ghrk::e3a_7 * A[7].y + ghrk::e3a_8 * A[8].y + ghrk::e3a_9 * A[9].y + ghrk::e3a_10 * A[10].y + ghrk::e3a_11 * A[11].y)) * isk;
    err3 += sqr * sqr;
    sqr = (ghrk::e5p * p0.y + h * (ghrk::e5a_1 * A[1].y + ghrk::e5a_4 * A[4].y + ghrk::e5a_5 * A[5].y + ghrk::e5a_6 * A[6].y + ghrk::e5a_7 * A[7].y + ghrk::e5a_8 * A[8].y + ghrk::e5a_9 * A[9].y + ghrk::e5a_10 * A[10].y + ghrk::e5a_11 * A[11].y)) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kp = ghdp::b1 * A[1].x + ghdp::b6 * A[6].x + ghdp::b7 * A[7].x + ghdp::b8 * A[8].x + ghdp::b9 * A[9].x + ghdp::b10 * A[10].x + ghdp::b11 * A[11].x + ghdp::b12 * A[12].x;
    pn.x = p0.x + h * kp;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(p0.x), fabs(pn.x)));
    double sqr = (kp - ghdp::bhh1 * A[1].x - ghdp::bhh2 * A[9].x - ghdp::bhh3 * A[12].x) * isk;
    err3 += sqr * sqr;
    sqr = (ghdp::er1 * A[1].x + ghdp::er6 * A[6].x + ghdp::er7 * A[7].x + ghdp::er8 * A[8].x + ghdp::er9 * A[9].x + ghdp::er10 * A[10].x + ghdp::er11 * A[11].x + ghdp::er12 * A[12].x) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kp = ghdp::b1 * A[1].y + ghdp::b6 * A[6].y + ghdp::b7 * A[7].y + ghdp::b8 * A[8].y + ghdp::b9 * A[9].y + ghdp::b10 * A[10].y + ghdp::b11 * A[11].y + ghdp::b12 * A[12].y;
    pn.y = p0.y + h * kp;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(p0.y), fabs(pn.y)));
    double sqr = (kp - ghdp::bhh1 * A[1].y - ghdp::bhh2 * A[9].y - ghdp::bhh3 * A[12].y) * isk;
    err3 += sqr * sqr;
    sqr = (ghdp::er1 * A[1].y + ghdp::er6 * A[6].y + ghdp::er7 * A[7].y + ghdp::er8 * A[8].y + ghdp::er9 * A[9].y + ghdp::er10 * A[10].y + ghdp::er11 * A[11].y + ghdp::er12 * A[12].y) * isk;
    err5 += sqr * sqr;
  }
}

// Dense output (leung_dop853.c:430-474) of the accepted step at the time
// t_old + sd h (s1 = 1 - sd), after the dense stages 14-16: per component
// the coefficients rc1..rc8 of the continuous extension, then
// rc1 + sd (rc2 + s1 (rc3 + sd (rc4 + s1 (rc5 + sd (rc6 + s1 (rc7 + sd rc8)))))).
__device__ __forceinline__ void gh_rkn_dense_eval(double h, double sd, double s1, const GhVec2& q0,
                                                 const GhVec2& p0, const GhVec2& qn,
                                                 const GhVec2& pn, const GhVec2* A, double* yo) {
  {
    const double ydiff = qn.x - q0.x;
    const double bspl = h * p0.x - ydiff;
    const double r0 = q0.x, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * pn.x - bspl;
    const double r4 = h * (ghrk::dp4 * p0.x + ghrk::dn4 * pn.x + h * (ghrk::da4_1 * A[1].x + ghrk::da4_4 * A[4].x + ghrk::da4_5 * A[5].x + ghrk::da4_6 * A[6].x + ghrk::da4_7 * A[7].x + ghrk::da4_8 * A[8].x + ghrk::da4_9 * A[9].x + ghrk::da4_10 * A[10].x + ghrk::da4_11 * A[11].x + ghrk::da4_12 * A[12].x + ghrk::da4_13 * A[13].x + ghrk::da4_14 * A[14].x + ghrk::da4_15 * A[15].x));
    const double r5 = h * (ghrk::dp5 * p0.x + ghrk::dn5 * pn.x + h * (ghrk::da5_1 * A[1].x + ghrk::da5_4 * A[4].x + ghrk::da5_5 * A[5].x + ghrk::da5_6 * A[6].x + ghrk::da5_7 * A[7].x + ghrk::da5_8 * A[8].x + ghrk::da5_9 * A[9].x + ghrk::da5_10 * A[10].x + ghrk::da5_11 * A[11].x + ghrk::da5_12 * A[12].x + ghrk::da5_13 * A[13].x + ghrk::da5_14 * A[14].x + ghrk::da5_15 * A[15].x));
    const double r6 = h * (ghrk::dp6 * p0.x + ghrk::dn6 * pn.x + h * (ghrk::da6_1 * A[1].x + ghrk::da6_4 * A[4].x + ghrk::da6_5 * A[5].x + ghrk::da6_6 * A[6].x + ghrk::da6_7 * A[7].x + ghrk::da6_8 * A[8].x + ghrk::da6_9 * A[9].x + ghrk::da6_10 * A[10].x + ghrk::da6_11 * A[11].x + ghrk::da6_12 * A[12].x + ghrk::da6_13 * A[13].x + ghrk::da6_14 * A[14].x + ghrk::da6_15 * A[15].x));
    const double r7 = h * (ghrk::dp7 * p0.x + ghrk::dn7 * pn.x + h * (ghrk::da7_1 * A[1].x + ghrk::da7_4 * A[4].x + ghrk::da7_5 * A[5].x + ghrk::da7_6 * A[6].x + ghrk::da7_7 * A[7].x + ghrk::da7_8 * A[8].x + ghrk::da7_9 * A[9].x + ghrk::da7_10 * A[10].x + ghrk::da7_11 * A[11].x + ghrk::da7_12 * A[12].x + ghrk::da7_13 * A[13].x + ghrk::da7_14 * A[14].x + ghrk::da7_15 * A[15].x));
    yo[0] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = qn.y - q0.y;
    const double bspl = h * p0.y - ydiff;
    const double r0 = q0.y, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * pn.y - bspl;
    const double r4 = h * (ghrk::dp4 * p0.y + ghrk::dn4 * pn.y + h * (ghrk::da4_1 * A[1].y + ghrk::da4_4 * A[4].y + ghrk::da4_5 * A[5].y + ghrk::da4_6 * A[6].y + ghrk::da4_7 * A[7].y + ghrk::da4_8 * A[8].y + ghrk::da4_9 * A[9].y + ghrk::da4_10 * A[10].y + ghrk::da4_11 * A[11].y + ghrk::da4_12 * A[12].y + ghrk::da4_13 * A[13].y + ghrk::da4_14 * A[14].y + ghrk::da4_15 * A[15].y));
    const double r5 = h * (ghrk::dp5 * p0.y + ghrk::dn5 * pn.y + h * (ghrk::da5_1 * A[1].y + ghrk::da5_4 * A[4].y + ghrk::da5_5 * A[5].y + ghrk::da5_6 * A[6].y + ghrk::da5_7 * A[7].y + ghrk::da5_8 * A[8].y + ghrk::da5_9 * A[9].y + ghrk::da5_10 * A[10].y + ghrk::da5_11 * A[11].y + ghrk::da5_12 * A[12].y + ghrk::da5_13 * A[13].y + ghrk::da5_14 * A[14].y + ghrk::da5_15 * A[15].y));
    const double r6 = h * (ghrk::dp6 * p0.y + ghrk::dn6 * pn.y + h * (ghrk::da6_1 * A[1].y + ghrk::da6_4 * A[4].y + ghrk::da6_5 * A[5].y + ghrk::da6_6 * A[6].y + ghrk::da6_7 * A[7].y + ghrk::da6_8 * A[8].y + ghrk::da6_9 * A[9].y + ghrk::da6_10 * A[10].y + ghrk::da6_11 * A[11].y + ghrk::da6_12 * A[12].y + ghrk::da6_13 * A[13].y + ghrk::da6_14 * A[14].y + ghrk::da6_15 * A[15].y));
    const double r7 = h * (ghrk::dp7 * p0.y + ghrk::dn7 * pn.y + h * (ghrk::da7_1 * A[1].y + ghrk::da7_4 * A[4].y + ghrk::da7_5 * A[5].y + ghrk::da7_6 * A[6].y + ghrk::da7_7 * A[7].y + ghrk::da7_8 * A[8].y + ghrk::da7_9 * A[9].y + ghrk::da7_10 * A[10].y + ghrk::da7_11 * A[11].y + ghrk::da7_12 * A[12].y + ghrk::da7_13 * A[13].y + ghrk::da7_14 * A[14].y + ghrk::da7_15 * A[15].y));
    yo[1] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = pn.x - p0.x;
    const double bspl = h * A[1].x - ydiff;
    const double r0 = p0.x, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * A[13].x - bspl;
    const double r4 = h * (ghdp::d41 * A[1].x + ghdp::d46 * A[6].x + ghdp::d47 * A[7].x + ghdp::d48 * A[8].x + ghdp::d49 * A[9].x + ghdp::d410 * A[10].x + ghdp::d411 * A[11].x + ghdp::d412 * A[12].x + ghdp::d413 * A[13].x + ghdp::d414 * A[14].x + ghdp::d415 * A[15].x + ghdp::d416 * A[16].x);
    const double r5 = h * (ghdp::d51 * A[1].x + ghdp::d56 * A[6].x + ghdp::d57 * A[7].x + ghdp::d58 * A[8].x + ghdp::d59 * A[9].x + ghdp::d510 * A[10].x + ghdp::d511 * A[11].x + ghdp::d512 * A[12].x + ghdp::d513 * A[13].x + ghdp::d514 * A[14].x + ghdp::d515 * A[15].x + ghdp::d516 * A[16].x);
    const double r6 = h * (ghdp::d61 * A[1].x + ghdp::d66 * A[6].x + ghdp::d67 * A[7].x + ghdp::d68 * A[8].x + ghdp::d69 * A[9].x + ghdp::d610 * A[10].x + ghdp::d611 * A[11].x + ghdp::d612 * A[12].x + ghdp::d613 * A[13].x + ghdp::d614 * A[14].x + ghdp::d615 * A[15].x + ghdp::d616 * A[16].x);
    const double r7 = h * (ghdp::d71 * A[1].x + ghdp::d76 * A[6].x + ghdp::d77 * A[7].x + ghdp::d78 * A[8].x + ghdp::d79 * A[9].x + ghdp::d710 * A[10].x + ghdp::d711 * A[11].x + ghdp::d712 * A[12].x + ghdp::d713 * A[13].x + ghdp::d714 * A[14].x + ghdp::d715 * A[15].x + ghdp::d716 * A[16].x);
    yo[2] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = pn.y - p0.y;
    const double bspl = h * A[1].y - ydiff;
    const double r0 = p0.y, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * A[13].y - bspl;
    const double r4 = h * (ghdp::d41 * A[1].y + ghdp::d46 * A[6].y + ghdp::d47 * A[7].y + ghdp::d48 * A[8].y + ghdp::d49 * A[9].y + ghdp::d410 * A[10].y + ghdp::d411 * A[11].y + ghdp::d412 * A[12].y + ghdp::d413 * A[13].y + ghdp::d414 * A[14].y + ghdp::d415 * A[15].y + ghdp::d416 * A[16].y);
    const double r5 = h * (ghdp::d51 * A[1].y + ghdp::d56 * A[6].y + ghdp::d57 * A[7].y + ghdp::d58 * A[8].y + ghdp::d59 * A[9].y + ghdp::d510 * A[10].y + ghdp::d511 * A[11].y + ghdp::d512 * A[12].y + ghdp::d513 * A[13].y + ghdp::d514 * A[14].y + ghdp::d515 * A[15].y + ghdp::d516 * A[16].y);
    const double r6 = h * (ghdp::d61 * A[1].y + ghdp::d66 * A[6].y + ghdp::d67 * A[7].y + ghdp::d68 * A[8].y + ghdp::d69 * A[9].y + ghdp::d610 * A[10].y + ghdp::d611 * A[11].y + ghdp::d612 * A[12].y + ghdp::d613 * A[13].y + ghdp::d614 * A[14].y + ghdp::d615 * A[15].y + ghdp::d616 * A[16].y);
    const double r7 = h * (ghdp::d71 * A[1].y + ghdp::d76 * A[6].y + ghdp::d77 * A[7].y + ghdp::d78 * A[8].y + ghdp::d79 * A[9].y + ghdp::d710 * A[10].y + ghdp::d711 * A[11].y + ghdp::d712 * A[12].y + ghdp::d713 * A[13].y + ghdp::d714 * A[14].y + ghdp::d715 * A[15].y + ghdp::d716 * A[16].y);
    yo[3] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
}
// position argument of stage 2
__device__ __forceinline__ GhVec6 gh_rkn_q2(double h, const GhVec6& q0, const GhVec6& p0,
                                             const GhVec6& pn, const GhVec6* A) {
  (void)pn;
  GhVec6 Q;
  Q.x = q0.x + h * (ghrk::qp2 * p0.x);
  Q.y = q0.y + h * (ghrk::qp2 * p0.y);
  Q.z = q0.z + h * (ghrk::qp2 * p0.z);
  Q.dx = q0.dx + h * (ghrk::qp2 * p0.dx);
  Q.dy = q0.dy + h * (ghrk::qp2 * p0.dy);
  Q.dz = q0.dz + h * (ghrk::qp2 * p0.dz);
  return Q;
}

// position argument of stage 3
__device__ __forceinline__ GhVec6 gh_rkn_q3(double h, const GhVec6& q0, const GhVec6& p0,
                                             const GhVec6& pn, const GhVec6* A) {
  (void)pn;
  GhVec6 Q;
  Q.x = q0.x + h * (ghrk::qp3 * p0.x + h * (ghrk::qa3_1 * A[1].x));
  Q.y = q0.y + h * (ghrk::qp3 * p0.y + h * (ghrk::qa3_1 * A[1].y));
  Q.z = q0.z + h * (ghrk::qp3 * p0.z + h * (ghrk::qa3_1 * A[1].z));
  Q.dx = q0.dx + h * (ghrk::qp3 * p0.dx + h * (ghrk::qa3_1 * A[1].dx));
  Q.dy = q0.dy + h * (ghrk::qp3 * p0.dy + h * (ghrk::qa3_1 * A[1].dy));
  Q.dz = q0.dz + h * (ghrk::qp3 * p0.dz + h * (ghrk::qa3_1 * A[1].dz));
  return Q;
}

// position argument of stage 4
__device__ __forceinline__ GhVec6 gh_rkn_q4(double h, const GhVec6& q0, const GhVec6& p0,
                                             const GhVec6& pn, const GhVec6* A) {
  (void)pn;
  GhVec6 Q;
  Q.x = q0.x + h * (ghrk::qp4 * p0.x + h * (ghrk::qa4_1 * A[1].x + ghrk::qa4_2 * A[2].x));
  Q.y = q0.y + h * (ghrk::qp4 * p0.y + h * (ghrk::qa4_1 * A[1].y + ghrk::qa4_2 * A[2].y));
  Q.z = q0.z + h * (ghrk::qp4 * p0.z + h * (ghrk::qa4_1 * A[1].z + ghrk::qa4_2 * A[2].z));
  Q.dx = q0.dx + h * (ghrk::qp4 * p0.dx + h * (ghrk::qa4_1 * A[1].dx + ghrk::qa4_2 * A[2].dx));
  Q.dy = q0.dy + h * (ghrk::qp4 * p0.dy + h * (ghrk::qa4_1 * A[1].dy + ghrk::qa4_2 * A[2].dy));
  Q.dz = q0.dz + h * (ghrk::qp4 * p0.dz + h * (ghrk::qa4_1 * A[1].dz + ghrk::qa4_2 * A[2].dz));
  return Q;
}

// position argument of stage 5
__device__ __forceinline__ GhVec6 gh_rkn_q5(double h, const GhVec6& q0, const GhVec6& p0,
                                             const GhVec6& pn, const GhVec6* A) {
  (void)pn;
  GhVec6 Q;
  Q.x = q0.x + h * (ghrk::qp5 * p0.x + h * (ghrk::qa5_1 * A[1].x + ghrk::qa5_2 * A[2].x + ghrk::qa5_3 * A[3].x));
  Q.y = q0.y + h * (ghrk::qp5 * p0.y + h * (ghrk::qa5_1 * A[1].y + ghrk::qa5_2 * A[2].y + ghrk::qa5_3 * A[3].y));
  Q.z = q0.z + h * (ghrk::qp5 * p0.z + h * (ghrk::qa5_1 * A[1].z + ghrk::qa5_2 * A[2].z + ghrk::qa5_3 * A[3].z));
  Q.dx = q0.dx + h * (ghrk::qp5 * p0.dx + h * (ghrk::qa5_1 * A[1].dx + ghrk::qa5_2 * A[2].dx + ghrk::qa5_3 * A[3].dx));
  Q.dy = q0.dy + h * (ghrk::qp5 * p0.dy + h * (ghrk::qa5_1 * A[1].dy + ghrk::qa5_2 * A[2].dy + ghrk::qa5_3 * A[3].dy));
  Q.dz = q0.dz + h * (ghrk::qp5 * p0.dz + h * (ghrk::qa5_1 * A[1].dz + ghrk::qa5_2 * A[2].dz + ghrk::qa5_3 * A[3].dz));
  return Q;
}

// position argument of stage 6
__device__ __forceinline__ GhVec6 gh_rkn_q6(double h, const GhVec6& q0, const GhVec6& p0,
                                             const GhVec6& pn, const GhVec6* A) {
  (void)pn;
  GhVec6 Q;
  Q.x = q0.x + h * (ghrk::qp6 * p0.x + h * (ghrk::qa6_1 * A[1].x + ghrk::qa6_3 * A[3].x + ghrk::qa6_4 * A[4].x));
  Q.y = q0.y + h * (ghrk::qp6 * p0.y + h * (ghrk::qa6_1 * A[1].y + ghrk::qa6_3 * A[3].y + ghrk::qa6_4 * A[4].y));
  Q.z = q0.z + h * (ghrk::qp6 * p0.z + h * (ghrk::qa6_1 * A[1].z + ghrk::qa6_3 * A[3].z + ghrk::qa6_4 * A[4].z));
  Q.dx = q0.dx + h * (ghrk::qp6 * p0.dx + h * (ghrk::qa6_1 * A[1].dx + ghrk::qa6_3 * A[3].dx + ghrk::qa6_4 * A[4].dx));
  Q.dy = q0.dy + h * (ghrk::qp6 * p0.dy + h * (ghrk::qa6_1 * A[1].dy + ghrk::qa6_3 * A[3].dy + ghrk::qa6_4 * A[4].dy));
  Q.dz = q0.dz + h * (ghrk::qp6 * p0.dz + h * (ghrk::qa6_1 * A[1].dz + ghrk::qa6_3 * A[3].dz + ghrk::qa6_4 * A[4].dz));
  return Q;
}

// position argument of stage 7
__device__ __forceinline__ GhVec6 gh_rkn_q7(double h, const GhVec6& q0, const GhVec6& p0,
                                             const GhVec6& pn, const GhVec6* A) {
  (void)pn;
  GhVec6 Q;
  Q.x = q0.x + h * (ghrk::qp7 * p0.x + h * (ghrk::qa7_1 * A[1].x + ghrk::qa7_3 * A[3].x + ghrk::qa7_4 * A[4].x + ghrk::qa7_5 * A[5].x));
  Q.y = q0.y + h * (ghrk::qp7 * p0.y + h * (ghrk::qa7_1 * A[1].y + ghrk::qa7_3 * A[3].y + ghrk::qa7_4 * A[4].y + ghrk::qa7_5 * A[5].y));
  Q.z = q0.z + h * (ghrk::qp7 * p0.z + h * (ghrk::qa7_1 * A[1].z + ghrk::qa7_3 * A[3].z + ghrk::qa7_4 * A[4].z + ghrk::qa7_5 * A[5].z));
  Q.dx = q0.dx + h * (ghrk::qp7 * p0.dx + h * (ghrk::qa7_1 * A[1].dx + ghrk::qa7_3 * A[3].dx + ghrk::qa7_4 * A[4].dx + ghrk::qa7_5 * A[5].dx));
  Q.dy = q0.dy + h * (ghrk::qp7 * p0.dy + h * (ghrk::qa7_1 * A[1].dy + ghrk::qa7_3 * A[3].dy + ghrk::qa7_4 * A[4].dy + ghrk::qa7_5 * A[5].dy));
  Q.dz = q0.dz + h * (ghrk::qp7 * p0.dz + h * (ghrk::qa7_1 * A[1].dz + ghrk::qa7_3 * A[3].dz + ghrk::qa7_4 * A[4].dz + ghrk::qa7_5 * A[5].dz));
  return Q;
}

// position argument of stage 8
__device__ __forceinline__ GhVec6 gh_rkn_q8(double h, const GhVec6& q0, const GhVec6& p0,
                                             const GhVec6& pn, const GhVec6* A) {
  (void)pn;
  GhVec6 Q;
  Q.x = q0.x + h * (ghrk::qp8 * p0.x + h * (ghrk::qa8_1 * A[1].x + ghrk::qa8_3 * A[3].x + ghrk::qa8_4 * A[4].x + ghrk::qa8_5 * A[5].x + ghrk::qa8_6 * A[6].x));
  Q.y = q0.y + h * (ghrk::qp8 * p0.y + h * (ghrk::qa8_1 * A[1].y + ghrk::qa8_3 * A[3].y + ghrk::qa8_4 * A[4].y + ghrk::qa8_5 * A[5].y + ghrk::qa8_6 * A[6].y));
  Q.z = q0.z + h * (ghrk::qp8 * p0.z + h * (ghrk::qa8_1 * A[1].z + ghrk::qa8_3 * A[3].z + ghrk::qa8_4 * A[4].z + ghrk::qa8_5 * A[5].z + ghrk::qa8_6 * A[6].z));
  Q.dx = q0.dx + h * (ghrk::qp8 * p0.dx + h * (ghrk::qa8_1 * A[1].dx + ghrk::qa8_3 * A[3].dx + ghrk::qa8_4 * A[4].dx + ghrk::qa8_5 * A[5].dx + ghrk::qa8_6 * A[6].dx));
  Q.dy = q0.dy + h * (ghrk::qp8 * p0.dy + h * (ghrk::qa8_1 * A[1].dy + ghrk::qa8_3 * A[3].dy + ghrk::qa8_4 * A[4].dy + ghrk::qa8_5 * A[5].dy + ghrk::qa8_6 * A[6].dy));
  Q.dz = q0.dz + h * (ghrk::qp8 * p0.dz + h * (ghrk::qa8_1 * A[1].dz + ghrk::qa8_3 * A[3].dz + ghrk::qa8_4 * A[4].dz + ghrk::qa8_5 * A[5].dz + ghrk::qa8_6 * A[6].dz));
  return Q;
}

// position argument of stage 9
__device__ __forceinline__ GhVec6 gh_rkn_q9(double h, const GhVec6& q0, const GhVec6& p0,
                                             const GhVec6& pn, const GhVec6* A) {
  (void)pn;
  GhVec6 Q;
  Q.x = q0.x + h * (ghrk::qp9 * p0.x + h * (ghrk::qa9_1 * A[1].x + ghrk::qa9_3 * A[3].x + ghrk::qa9_4 * A[4].x + ghrk::qa9_5 * A[5].x + ghrk::qa9_6 * A[6].x + ghrk::qa9_7 * A[7].x));
  Q.y = q0.y + h * (ghrk::qp9 * p0.y + h * (ghrk::qa9_1 * A[1].y + ghrk::qa9_3 * A[3].y + ghrk::qa9_4 * A[4].y + ghrk::qa9_5 * A[5].y + ghrk::qa9_6 * A[6].y + ghrk::qa9_7 * A[7].y));
  Q.z = q0.z + h * (ghrk::qp9 * p0.z + h * (ghrk::qa9_1 * A[1].z + ghrk::qa9_3 * A[3].z + ghrk::qa9_4 * A[4].z + ghrk::qa9_5 * A[5].z + ghrk::qa9_6 * A[6].z + ghrk::qa9_7 * A[7].z));
  Q.dx = q0.dx + h * (ghrk::qp9 * p0.dx + h * (ghrk::qa9_1 * A[1].dx + ghrk::qa9_3 * A[3].dx + ghrk::qa9_4 * A[4].dx + ghrk::qa9_5 * A[5].dx + ghrk::qa9_6 * A[6].dx + ghrk::qa9_7 * A[7].dx));
  Q.dy = q0.dy + h * (ghrk::qp9 * p0.dy + h * (ghrk::qa9_1 * A[1].dy + ghrk::qa9_3 * A[3].dy + ghrk::qa9_4 * A[4].dy + ghrk::qa9_5 * A[5].dy + ghrk::qa9_6 * A[6].dy + ghrk::qa9_7 * A[7].dy));
  Q.dz = q0.dz + h * (ghrk::qp9 * p0.dz + h * (ghrk::qa9_1 * A[1].dz + ghrk::qa9_3 * A[3].dz + ghrk::qa9_4 * A[4].dz + ghrk::qa9_5 * A[5].dz + ghrk::qa9_6 * A[6].dz + ghrk::qa9_7 * A[7].dz));
  return Q;
}

// position argument of stage 10
__device__ __forceinline__ GhVec6 gh_rkn_q10(double h, const GhVec6& q0, const GhVec6& p0,
                                             const GhVec6& pn, const GhVec6* A) {
  (void)pn;
  GhVec6 Q;
  Q.x = q0.x + h * (ghrk::qp10 * p0.x + h * (ghrk::qa10_1 * A[1].x + ghrk::qa10_3 * A[3].x + ghrk::qa10_4 * A[4].x + ghrk::qa10_5 * A[5].x + ghrk::qa10_6 * A[6].x + ghrk::qa10_7 * A[7].x + ghrk::qa10_8 * A[8].x));
  Q.y = q0.y + h * (ghrk::qp10 * p0.y + h * (ghrk::qa10_1 * A[1].y + ghrk::qa10_3 * A[3].y + ghrk::qa10_4 * A[4].y + ghrk::qa10_5 * A[5].y + ghrk::qa10_6 * A[6].y + ghrk::qa10_7 * A[7].y + ghrk::qa10_8 * A[8].y));
  Q.z = q0.z + h * (ghrk::qp10 * p0.z + h * (ghrk::qa10_1 * A[1].z + ghrk::qa10_3 * A[3].z + ghrk::qa10_4 * A[4].z + ghrk::qa10_5 * A[5].z + ghrk::qa10_6 * A[6].z + ghrk::qa10_7 * A[7].z + ghrk::qa10_8 * A[8].z));
  Q.dx = q0.dx + h * (ghrk::qp10 * p0.dx + h * (ghrk::qa10_1 * A[1].dx + ghrk::qa10_3 * A[3].dx + ghrk::qa10_4 * A[4].dx + ghrk::qa10_5 * A[5].dx + ghrk::qa10_6 * A[6].dx + ghrk::qa10_7 * A[7].dx + ghrk::qa10_8 * A[8].dx));
  Q.dy = q0.dy + h * (ghrk::qp10 * p0.dy + h * (ghrk::qa10_1 * A[1].dy + ghrk::qa10_3 * A[3].dy + ghrk::qa10_4 * A[4].dy + ghrk::qa10_5 * A[5].dy + ghrk::qa10_6 * A[6].dy + ghrk::qa10_7 * A[7].dy + ghrk::qa10_8 * A[8].dy));
  Q.dz = q0.dz + h * (ghrk::qp10 * p0.dz + h * (ghrk::qa10_1 * A[1].dz + ghrk::qa10_3 * A[3].dz + ghrk::qa10_4 * A[4].dz + ghrk::qa10_5 * A[5].dz + ghrk::qa10_6 * A[6].dz + ghrk::qa10_7 * A[7].dz + ghrk::qa10_8 * A[8].dz));
  return Q;
}

// position argument of stage 11
__device__ __forceinline__ GhVec6 gh_rkn_q11(double h, const GhVec6& q0, const GhVec6& p0,
                                             const GhVec6& pn, const GhVec6* A) {
  (void)pn;
  GhVec6 Q;
  Q.x = q0.x + h * (ghrk::qp11 * p0.x + h * (ghrk::qa11_1 * A[1].x + ghrk::qa11_3 * A[3].x + ghrk::qa11_4 * A[4].x + ghrk::qa11_5 * A[5].x + ghrk::qa11_6 * A[6].x + ghrk::qa11_7 * A[7].x + ghrk::qa11_8 * A[8].x + ghrk::qa11_9 * A[9].x));
  Q.y = q0.y + h * (ghrk::qp11 * p0.y + h * (ghrk::qa11_1 * A[1].y + ghrk::qa11_3 * A[3].y + ghrk::qa11_4 * A[4].y + ghrk::qa11_5 * A[5].y + ghrk::qa11_6 * A[6].y + ghrk::qa11_7 * A[7].y + ghrk::qa11_8 * A[8].y + ghrk::qa11_9 * A[9].y));
  Q.z = q0.z + h * (ghrk::qp11 * p0.z + h * (ghrk::qa11_1 * A[1].z + ghrk::qa11_3 * A[3].z + ghrk::qa11_4 * A[4].z + ghrk::qa11_5 * A[5].z + ghrk::qa11_6 * A[6].z + ghrk::qa11_7 * A[7].z + ghrk::qa11_8 * A[8].z + ghrk::qa11_9 * A[9].z));
  Q.dx = q0.dx + h * (ghrk::qp11 * p0.dx + h * (ghrk::qa11_1 * A[1].dx + ghrk::qa11_3 * A[3].dx + ghrk::qa11_4 * A[4].dx + ghrk::qa11_5 * A[5].dx + ghrk::qa11_6 * A[6].dx + ghrk::qa11_7 * A[7].dx + ghrk::qa11_8 * A[8].dx + ghrk::qa11_9 * A[9].dx));
  Q.dy = q0.dy + h * (ghrk::qp11 * p0.dy + h * (ghrk::qa11_1 * A[1].dy + ghrk::qa11_3 * A[3].dy + ghrk::qa11_4 * A[4].dy + ghrk::qa11_5 * A[5].dy + ghrk::qa11_6 * A[6].dy + ghrk::qa11_7 * A[7].dy + ghrk::qa11_8 * A[8].dy + ghrk::qa11_9 * A[9].dy));
  Q.dz = q0.dz + h * (ghrk::qp11 * p0.dz + h * (ghrk::qa11_1 * A[1].dz + ghrk::qa11_3 * A[3].dz + ghrk::qa11_4 * A[4].dz + ghrk::qa11_5 * A[5].dz + ghrk::qa11_6 * A[6].dz + ghrk::qa11_7 * A[7].dz + ghrk::qa11_8 * A[8].dz + ghrk::qa11_9 * A[9].dz));
  return Q;
}

// position argument of stage 12
__device__ __forceinline__ GhVec6 gh_rkn_q12(double h, const GhVec6& q0, const GhVec6& p0,
                                             const GhVec6& pn, const GhVec6* A) {
  (void)pn;
  GhVec6 Q;
  Q.x = q0.x + h * (ghrk::qp12 * p0.x + h * (ghrk::qa12_1 * A[1].x + ghrk::qa12_3 * A[3].x + ghrk::qa12_4 * A[4].x + ghrk::qa12_5 * A[5].x + ghrk::qa12_6 * A[6].x + ghrk::qa12_7 * A[7].x + ghrk::qa12_8 * A[8].x + ghrk::qa12_9 * A[9].x + ghrk::qa12_10 * A[10].x));
  Q.y = q0.y + h * (ghrk::qp12 * p0.y + h * (ghrk::qa12_1 * A[1].y + ghrk::qa12_3 * A[3].y + ghrk::qa12_4 * A[4].y + ghrk::qa12_5 * A[5].y + ghrk::qa12_6 * A[6].y + ghrk::qa12_7 * A[7].y + ghrk::qa12_8 * A[8].y + ghrk::qa12_9 * A[9].y + ghrk::qa12_10 * A[10].y));
  Q.z = q0.z + h * (ghrk::qp12 * p0.z + h * (ghrk::qa12_1 * A[1].z + ghrk::qa12_3 * A[3].z + ghrk::qa12_4 * A[4].z + ghrk::qa12_5 * A[5].z + ghrk::qa12_6 * A[6].z + ghrk::qa12_7 * A[7].z + ghrk::qa12_8 * A[8].z + ghrk::qa12_9 * A[9].z + ghrk::qa12_10 * A[10].z));
  Q.dx = q0.dx + h * (ghrk::qp12 * p0.dx + h * (ghrk::qa12_1 * A[1].dx + ghrk::qa12_3 * A[3].dx + ghrk::qa12_4 * A[4].dx + ghrk::qa12_5 * A[5].dx + ghrk::qa12_6 * A[6].dx + ghrk::qa12_7 * A[7].dx + ghrk::qa12_8 * A[8].dx + ghrk::qa12_9 * A[9].dx + ghrk::qa12_10 * A[10].dx));
  Q.dy = q0.dy + h * (ghrk::qp12 * p0.dy + h * (ghrk::qa12_1 * A[1].dy + ghrk::qa12_3 * A[3].dy + ghrk::qa12_4 * A[4].dy + ghrk::qa12_5 * A[5].dy + ghrk::qa12_6 * A[6].dy + ghrk::qa12_7 * A[7].dy + ghrk::qa12_8 * A[8].dy + ghrk::qa12_9 * A[9].dy + ghrk::qa12_10 * A[10].dy));
  Q.dz = q0.dz + h * (ghrk::qp12 * p0.dz + h * (ghrk::qa12_1 * A[1].dz + ghrk::qa12_3 * A[3].dz + ghrk::qa12_4 * A[4].dz + ghrk::qa12_5 * A[5].dz + ghrk::qa12_6 * A[6].dz + ghrk::qa12_7 * A[7].dz + ghrk::qa12_8 * A[8].dz + ghrk::qa12_9 * A[9].dz + ghrk::qa12_10 * A[10].dz));
  return Q;
}

// position argument of stage 14
__device__ __forceinline__ GhVec6 gh_rkn_q14(double h, const GhVec6& q0, const GhVec6& p0,
                                             const GhVec6& pn, const GhVec6* A) {
  (void)pn;
  GhVec6 Q;
  Q.x = q0.x + h * (ghrk::qp14 * p0.x + ghrk::qn14 * pn.x + h * (ghrk::qa14_1 * A[1].x + ghrk::qa14_4 * A[4].x + ghrk::qa14_5 * A[5].x + ghrk::qa14_6 * A[6].x + ghrk::qa14_7 * A[7].x + ghrk::qa14_8 * A[8].x + ghrk::qa14_9 * A[9].x + ghrk::qa14_10 * A[10].x + ghrk::qa14_11 * A[11].x));
  Q.y = q0.y + h * (ghrk::qp14 * p0.y + ghrk::qn14 * pn.y + h * (ghrk::qa14_1 * A[1].y + ghrk::qa14_4 * A[4].y + ghrk::qa14_5 * A[5].y + ghrk::qa14_6 * A[6].y + ghrk::qa14_7 * A[7].y + ghrk::qa14_8 * A[8].y + ghrk::qa14_9 * A[9].y + ghrk::qa14_10 * A[10].y + ghrk::qa14_11 * A[11].y));
  Q.z = q0.z + h * (ghrk::qp14 * p0.z + ghrk::qn14 * pn.z + h * (ghrk::qa14_1 * A[1].z + ghrk::qa14_4 * A[4].z + ghrk::qa14_5 * A[5].z + ghrk::qa14_6 * A[6].z + ghrk::qa14_7 * A[7].z + ghrk::qa14_8 * A[8].z + ghrk::qa14_9 * A[9].z + ghrk::qa14_10 * A[10].z + ghrk::qa14_11 * A[11].z));
  Q.dx = q0.dx + h * (ghrk::qp14 * p0.dx + ghrk::qn14 * pn.dx + h * (ghrk::qa14_1 * A[1].dx + ghrk::qa14_4 * A[4].dx + ghrk::qa14_5 * A[5].dx + ghrk::qa14_6 * A[6].dx + ghrk::qa14_7 * A[7].dx + ghrk::qa14_8 * A[8].dx + ghrk::qa14_9 * A[9].dx + ghrk::qa14_10 * A[10].dx + ghrk::qa14_11 * A[11].dx));
  Q.dy = q0.dy + h * (ghrk::qp14 * p0.dy + ghrk::qn14 * pn.dy + h * (ghrk::qa14_1 * A[1].dy + ghrk::qa14_4 * A[4].dy + ghrk::qa14_5 * A[5].dy + ghrk::qa14_6 * A[6].dy + ghrk::qa14_7 * A[7].dy + ghrk::qa14_8 * A[8].dy + ghrk::qa14_9 * A[9].dy + ghrk::qa14_10 * A[10].dy + ghrk::qa14_11 * A[11].dy));
  Q.dz = q0.dz + h * (ghrk::qp14 * p0.dz + ghrk::qn14 * pn.dz + h * (ghrk::qa14_1 * A[1].dz + ghrk::qa14_4 * A[4].dz + ghrk::qa14_5 * A[5].dz + ghrk::qa14_6 * A[6].dz + ghrk::qa14_7 * A[7].dz + ghrk::qa14_8 * A[8].dz + ghrk::qa14_9 * A[9].dz + ghrk::qa14_10 * A[10].dz + ghrk::qa14_11 * A[11].dz));
  return Q;
}

// position argument of stage 15
__device__ __forceinline__ GhVec6 gh_rkn_q15(double h, const GhVec6& q0, const GhVec6& p0,
                                             const GhVec6& pn, const GhVec6* A) {
  (void)pn;
  GhVec6 Q;
  Q.x = q0.x + h * (ghrk::qp15 * p0.x + ghrk::qn15 * pn.x + h * (ghrk::qa15_1 * A[1].x + ghrk::qa15_4 * A[4].x + ghrk::qa15_5 * A[5].x + ghrk::qa15_6 * A[6].x + ghrk::qa15_7 * A[7].x + ghrk::qa15_8 * A[8].x + ghrk::qa15_9 * A[9].x + ghrk::qa15_10 * A[10].x + ghrk::qa15_11 * A[11].x + ghrk::qa15_12 * A[12].x + ghrk::qa15_13 * A[13].x));
  Q.y = q0.y + h * (ghrk::qp15 * p0.y + ghrk::qn15 * pn.y + h * (ghrk::qa15_1 * A[1].y + ghrk::qa15_4 * A[4].y + ghrk::qa15_5 * A[5].y + ghrk::qa15_6 * A[6].y + ghrk::qa15_7 * A[7].y + ghrk::qa15_8 * A[8].y + ghrk::qa15_9 * A[9].y + ghrk::qa15_10 * A[10].y + ghrk::qa15_11 * A[11].y + ghrk::qa15_12 * A[12].y + ghrk::qa15_13 * A[13].y));
  Q.z = q0.z + h * (ghrk::qp15 * p0.z + ghrk::qn15 * pn.z + h * (ghrk::qa15_1 * A[1].z + ghrk::qa15_4 * A[4].z + ghrk::qa15_5 * A[5].z + ghrk::qa15_6 * A[6].z + ghrk::qa15_7 * A[7].z + ghrk::qa15_8 * A[8].z + ghrk::qa15_9 * A[9].z + ghrk::qa15_10 * A[10].z + ghrk::qa15_11 * A[11].z + ghrk::qa15_12 * A[12].z + ghrk::qa15_13 * A[13].z));
  Q.dx = q0.dx + h * (ghrk::qp15 * p0.dx + ghrk::qn15 * pn.dx + h * (ghrk::qa15_1 * A[1].dx + ghrk::qa15_4 * A[4].dx + ghrk::qa15_5 * A[5].dx + ghrk::qa15_6 * A[6].dx + ghrk::qa15_7 * A[7].dx + ghrk::qa15_8 * A[8].dx + ghrk::qa15_9 * A[9].dx + ghrk::qa15_10 * A[10].dx + ghrk::qa15_11 * A[11].dx + ghrk::qa15_12 * A[12].dx + ghrk::qa15_13 * A[13].dx));
  Q.dy = q0.dy + h * (ghrk::qp15 * p0.dy + ghrk::qn15 * pn.dy + h * (ghrk::qa15_1 * A[1].dy + ghrk::qa15_4 * A[4].dy + ghrk::qa15_5 * A[5].dy + ghrk::qa15_6 * A[6].dy + ghrk::qa15_7 * A[7].dy + ghrk::qa15_8 * A[8].dy + ghrk::qa15_9 * A[9].dy + ghrk::qa15_10 * A[10].dy + ghrk::qa15_11 * A[11].dy + ghrk::qa15_12 * A[12].dy + ghrk::qa15_13 * A[13].dy));
  Q.dz = q0.dz + h * (ghrk::qp15 * p0.dz + ghrk::qn15 * pn.dz + h * (ghrk::qa15_1 * A[1].dz + ghrk::qa15_4 * A[4].dz + ghrk::qa15_5 * A[5].dz + ghrk::qa15_6 * A[6].dz + ghrk::qa15_7 * A[7].dz + ghrk::qa15_8 * A[8].dz + ghrk::qa15_9 * A[9].dz + ghrk::qa15_10 * A[10].dz + ghrk::qa15_11 * A[11].dz + ghrk::qa15_12 * A[12].dz + ghrk::qa15_13 * A[13].dz));
  return Q;
}

// position argument of stage 16
__device__ __forceinline__ GhVec6 gh_rkn_q16(double h, const GhVec6& q0, const GhVec6& p0,
                                             const GhVec6& pn, const GhVec6* A) {
  (void)pn;
  GhVec6 Q;
  Q.x = q0.x + h * (ghrk::qp16 * p0.x + ghrk::qn16 * pn.x + h * (ghrk::qa16_1 * A[1].x + ghrk::qa16_4 * A[4].x + ghrk::qa16_5 * A[5].x + ghrk::qa16_6 * A[6].x + ghrk::qa16_7 * A[7].x + ghrk::qa16_8 * A[8].x + ghrk::qa16_9 * A[9].x + ghrk::qa16_10 * A[10].x + ghrk::qa16_11 * A[11].x + ghrk::qa16_12 * A[12].x + ghrk::qa16_13 * A[13].x + ghrk::qa16_14 * A[14].x));
  Q.y = q0.y + h * (ghrk::qp16 * p0.y + ghrk::qn16 * pn.y + h * (ghrk::qa16_1 * A[1].y + ghrk::qa16_4 * A[4].y + ghrk::qa16_5 * A[5].y + ghrk::qa16_6 * A[6].y + ghrk::qa16_7 * A[7].y + ghrk::qa16_8 * A[8].y + ghrk::qa16_9 * A[9].y + ghrk::qa16_10 * A[10].y + ghrk::qa16_11 * A[11].y + ghrk::qa16_12 * A[12].y + ghrk::qa16_13 * A[13].y + ghrk::qa16_14 * A[14].y));
  Q.z = q0.z + h * (ghrk::qp16 * p0.z + ghrk::qn16 * pn.z + h * (ghrk::qa16_1 * A[1].z + ghrk::qa16_4 * A[4].z + ghrk::qa16_5 * A[5].z + ghrk::qa16_6 * A[6].z + ghrk::qa16_7 * A[7].z + ghrk::qa16_8 * A[8].z + ghrk::qa16_9 * A[9].z + ghrk::qa16_10 * A[10].z + ghrk::qa16_11 * A[11].z + ghrk::qa16_12 * A[12].z + ghrk::qa16_13 * A[13].z + ghrk::qa16_14 * A[14].z));
  Q.dx = q0.dx + h * (ghrk::qp16 * p0.dx + ghrk::qn16 * pn.dx + h * (ghrk::qa16_1 * A[1].dx + ghrk::qa16_4 * A[4].dx + ghrk::qa16_5 * A[5].dx + ghrk::qa16_6 * A[6].dx + ghrk::qa16_7 * A[7].dx + ghrk::qa16_8 * A[8].dx + ghrk::qa16_9 * A[9].dx + ghrk::qa16_10 * A[10].dx + ghrk::qa16_11 * A[11].dx + ghrk::qa16_12 * A[12].dx + ghrk::qa16_13 * A[13].dx + ghrk::qa16_14 * A[14].dx));
  Q.dy = q0.dy + h * (ghrk::qp16 * p0.dy + ghrk::qn16 * pn.dy + h * (ghrk::qa16_1 * A[1].dy + ghrk::qa16_4 * A[4].dy + ghrk::qa16_5 * A[5].dy + ghrk::qa16_6 * A[6].dy + ghrk::qa16_7 * A[7].dy + ghrk::qa16_8 * A[8].dy + ghrk::qa16_9 * A[9].dy + ghrk::qa16_10 * A[10].dy + ghrk::qa16_11 * A[11].dy + ghrk::qa16_12 * A[12].dy + ghrk::qa16_13 * A[13].dy + ghrk::qa16_14 * A[14].dy));
  Q.dz = q0.dz + h * (ghrk::qp16 * p0.dz + ghrk::qn16 * pn.dz + h * (ghrk::qa16_1 * A[1].dz + ghrk::qa16_4 * A[4].dz + ghrk::qa16_5 * A[5].dz + ghrk::qa16_6 * A[6].dz + ghrk::qa16_7 * A[7].dz + ghrk::qa16_8 * A[8].dz + ghrk::qa16_9 * A[9].dz + ghrk::qa16_10 * A[10].dz + ghrk::qa16_11 * A[11].dz + ghrk::qa16_12 * A[12].dz + ghrk::qa16_13 * A[13].dz + ghrk::qa16_14 * A[14].dz));
  return Q;
}

// New state (q_new, p_new) and the two error sums of leung_dop853.c:376-404:
// per component, sk = atol + rtol max(|y0|, |y_new|), err3 += (e3/sk)^2,
// err5 += (e5/sk)^2, with e3 = sum (b - bhh) k, e5 = sum er k (one
// reciprocal of sk per component; the norm only steers the step size).
__device__ __forceinline__ void gh_rkn_solution(double h, const GhVec6& q0, const GhVec6& p0,
                                               const GhVec6* A, double atol, double rtol,
                                               GhVec6& qn, GhVec6& pn, double& err5,
                                               double& err3) {
  err5 = 0.0;
  err3 = 0.0;
  {
    const double kq = ghrk::bp * p0.x + h * (ghrk::ba_1 * A[1].x + ghrk::ba_4 * A[4].x + ghrk::ba_5 * A[5].x + ghrk::ba_6 * A[6].x + ghrk::ba_7 * A[7].x + ghrk::ba_8 * A[8].x + ghrk::ba_9 * A[9].x + ghrk::ba_10 * A[10].x + ghrk::ba_11 * A[11].x);
    qn.x = q0.x + h * kq;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(q0.x), fabs(qn.x)));
    double sqr = (ghrk::e3p * p0.x + h * (ghrk::e3a_1 * A[1].x + ghrk::e3a_4 * A[4].x + ghrk::e3a_5 * A[5].x + ghrk::e3a_6 * A[6].x + ghrk::e3a_7 * A[7].x + ghrk::e3a_8 * A[8].x + ghrk::e3a_9 * A[9].x + ghrk::e3a_10 * A[10].x + ghrk::e3a_11 * A[11].x)) * isk;
    err3 += sqr * sqr;
    sqr = (ghrk::e5p * p0.x + h * (ghrk::e5a_1 * A[1].x + ghrk::e5a_4 * A[4].x + ghrk::e5a_5 * A[5].x + ghrk::e5a_6 * A[6].x + ghrk::e5a_7 * A[7].x + ghrk::e5a_8 * A[8].x + ghrk::e5a_9 * A[9].x + ghrk::e5a_10 * A[10].x + ghrk::e5a_11 * A[11].x)) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kq = ghrk::bp * p0.y + h * (ghrk::ba_1 * A[1].y + ghrk::ba_4 * A[4].y + ghrk::ba_5 * A[5].y + ghrk::ba_6 * A[6].y + ghrk::ba_7 * A[7].y + ghrk::ba_8 * A[8].y + ghrk::ba_9 * A[9].y + ghrk::ba_10 * A[10].y + ghrk::ba_11 * A[11].y);
    qn.y = q0.y + h * kq;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(q0.y), fabs(qn.y)));
    double sqr = (ghrk::e3p * p0.y + h * (ghrk::e3a_1 * A[1].y + ghrk::e3a_4 * A[4].y + ghrk::e3a_5 * A[5].y + ghrk::e3a_6 * A[6].y + ghrk::e3a_7 * A[7].y + ghrk::e3a_8 * A[8].y + ghrk::e3a_9 * A[9].y + ghrk::e3a_10 * A[10].y + ghrk::e3a_11 * A[11].y)) * isk;
    err3 += sqr * sqr;
    sqr = (ghrk::e5p * p0.y + h * (ghrk::e5a_1 * A[1].y + ghrk::e5a_4 * A[4].y + ghrk::e5a_5 * A[5].y + ghrk::e5a_6 * A[6].y + ghrk::e5a_7 * A[7].y + ghrk::e5a_8 * A[8].y + ghrk::e5a_9 * A[9].y + ghrk::e5a_10 * A[10].y + ghrk::e5a_11 * A[11].y)) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kq = ghrk::bp * p0.z + h * (ghrk::ba_1 * A[1].z + ghrk::ba_4 * A[4].z + ghrk::ba_5 * A[5].z + ghrk::ba_6 * A[6].z + ghrk::ba_7 * A[7].z + ghrk::ba_8 * A[8].z + ghrk::ba_9 * A[9].z + ghrk::ba_10 * A[10].z + ghrk::ba_11 * A[11].z);
    qn.z = q0.z + h * kq;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(q0.z), fabs(qn.z)));
    double sqr = (ghrk::e3p * p0.z + h * (ghrk::e3a_1 * A[1].z + ghrk::e3a_4 * A[4].z + ghrk::e3a_5 * A[5].z + ghrk::e3a_6 * A[6].z + ghrk::e3a_7 * A[7].z + ghrk::e3a_8 * A[8].z + ghrk::e3a_9 * A[9].z + ghrk::e3a_10 * A[10].z + ghrk::e3a_11 * A[11].z)) * isk;
    err3 += sqr * sqr;
    sqr = (ghrk::e5p * p0.z + h * (ghrk::e5a_1 * A[1].z + ghrk::e5a_4 * A[4].z + ghrk::e5a_5 * A[5].z + ghrk::e5a_6 * A[6].z + ghrk::e5a_7 * A[7].z + ghrk::e5a_8 * A[8].z + ghrk::e5a_9 * A[9].z + ghrk::e5a_10 * A[10].z + ghrk::e5a_11 * A[11].z)) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kq = ghrk::bp * p0.dx + h * (ghrk::ba_1 * A[1].dx + ghrk::ba_4 * A[4].dx + ghrk::ba_5 * A[5].dx + ghrk::ba_6 * A[6].dx + ghrk::ba_7 * A[7].dx + ghrk::ba_8 * A[8].dx + ghrk::ba_9 * A[9].dx + ghrk::ba_10 * A[10].dx + ghrk::ba_11 * A[11].dx);
    qn.dx = q0.dx + h * kq;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(q0.dx), fabs(qn.dx)));
    double sqr = (ghrk::e3p * p0.dx + h * (ghrk::e3a_1 * A[1].dx + ghrk::e3a_4 * A[4].dx + ghrk::e3a_5 * A[5].dx + ghrk::e3a_6 * A[6].dx + ghrk::e3a_7 * A[7].dx + ghrk::e3a_8 * A[8].dx + ghrk::e3a_9 * A[9].dx + ghrk::e3a_10 * A[10].dx + ghrk::e3a_11 * A[11].dx)) * isk;
    err3 += sqr * sqr;
    sqr = (ghrk::e5p * p0.dx + h * (ghrk::e5a_1 * A[1].dx + ghrk::e5a_4 * A[4].dx + ghrk::e5a_5 * A[5].dx + ghrk::e5a_6 * A[6].dx + ghrk::e5a_7 * A[7].dx + ghrk::e5a_8 * A[8].dx + ghrk::e5a_9 * A[9].dx + ghrk::e5a_10 * A[10].dx + ghrk::e5a_11 * A[11].dx)) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kq = ghrk::bp * p0.dy + h * (ghrk::ba_1 * A[1].dy + ghrk::ba_4 * A[4].dy + ghrk::ba_5 * A[5].dy + ghrk::ba_6 * A[6].dy + ghrk::ba_7 * A[7].dy + ghrk::ba_8 * A[8].dy + ghrk::ba_9 * A[9].dy + ghrk::ba_10 * A[10].dy + ghrk::ba_11 * A[11].dy);
    qn.dy = q0.dy + h * kq;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(q0.dy), fabs(qn.dy)));
    double sqr = (ghrk::e3p * p0.dy + h * (ghrk::e3a_1 * A[1].dy + ghrk::e3a_4 * A[4].dy + ghrk::e3a_5 * A[5].dy + ghrk::e3a_6 * A[6].dy + ghrk::e3a_7 * A[7].dy + ghrk::e3a_8 * A[8].dy + ghrk::e3a_9 * A[9].dy + ghrk::e3a_10 * A[10].dy + ghrk::e3a_11 * A[11].dy)) * isk;
    err3 += sqr * sqr;
    sqr = (ghrk::e5p * p0.dy + h * (ghrk::e5a_1 * A[1].dy + ghrk::e5a_4 * A[4].dy + ghrk::e5a_5 * A[5].dy + ghrk::e5a_6 * A[6].dy + ghrk::e5a_7 * A[7].dy + ghrk::e5a_8 * A[8].dy + ghrk::e5a_9 * A[9].dy + ghrk::e5a_10 * A[10].dy + ghrk::e5a_11 * A[11].dy)) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kq = ghrk::bp * p0.dz + h * (ghrk::ba_1 * A[1].dz + ghrk::ba_4 * A[4].dz + ghrk::ba_5 * A[5].dz + ghrk::ba_6 * A[6].dz + ghrk::ba_7 * A[7].dz + ghrk::ba_8 * A[8].dz + ghrk::ba_9 * A[9].dz + ghrk::ba_10 * A[10].dz + ghrk::ba_11 * A[11].dz);
    qn.dz = q0.dz + h * kq;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(q0.dz), fabs(qn.dz)));
    double sqr = (ghrk::e3p * p0.dz + h * (ghrk::e3a_1 * A[1].dz + ghrk::e3a_4 * A[4].dz + ghrk::e3a_5 * A[5].dz + ghrk::e3a_6 * A[6].dz + ghrk::e3a_7 * A[7].dz + ghrk::e3a_8 * A[8].dz + ghrk::e3a_9 * A[9].dz + ghrk::e3a_10 * A[10].dz + ghrk::e3a_11 * A[11].dz)) * isk;
    err3 += sqr * sqr;
    sqr = (ghrk::e5p * p0.dz + h * (ghrk::e5a_1 * A[1].dz + ghrk::e5a_4 * A[4].dz + ghrk::e5a_5 * A[5].dz + ghrk::e5a_6 * A[6].dz + ghrk::e5a_7 * A[7].dz + ghrk::e5a_8 * A[8].dz + ghrk::e5a_9 * A[9].dz + ghrk::e5a_10 * A[10].dz + ghrk::e5a_11 * A[11].dz)) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kp = ghdp::b1 * A[1].x + ghdp::b6 * A[6].x + ghdp::b7 * A[7].x + ghdp::b8 * A[8].x + ghdp::b9 * A[9].x + ghdp::b10 * A[10].x + ghdp::b11 * A[11].x + ghdp::b12 * A[12].x;
    pn.x = p0.x + h * kp;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(p0.x), fabs(pn.x)));
    double sqr = (kp - ghdp::bhh1 * A[1].x - ghdp::bhh2 * A[9].x - ghdp::bhh3 * A[12].x) * isk;
    err3 += sqr * sqr;
    sqr = (ghdp::er1 * A[1].x + ghdp::er6 * A[6].x + ghdp::er7 * A[7].x + ghdp::er8 * A[8].x + ghdp::er9 * A[9].x + ghdp::er10 * A[10].x + ghdp::er11 * A[11].x + ghdp::er12 * A[12].x) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kp = ghdp::b1 * A[1].y + ghdp::b6 * A[6].y + ghdp::b7 * A[7].y + ghdp::b8 * A[8].y + ghdp::b9 * A[9].y + ghdp::b10 * A[10].y + ghdp::b11 * A[11].y + ghdp::b12 * A[12].y;
    pn.y = p0.y + h * kp;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(p0.y), fabs(pn.y)));
    double sqr = (kp - ghdp::bhh1 * A[1].y - ghdp::bhh2 * A[9].y - ghdp::bhh3 * A[12].y) * isk;
    err3 += sqr * sqr;
    sqr = (ghdp::er1 * A[1].y + ghdp::er6 * A[6].y + ghdp::er7 * A[7].y + ghdp::er8 * A[8].y + ghdp::er9 * A[9].y + ghdp::er10 * A[10].y + ghdp::er11 * A[11].y + ghdp::er12 * A[12].y) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kp = ghdp::b1 * A[1].z + ghdp::b6 * A[6].z + ghdp::b7 * A[7].z + ghdp::b8 * A[8].z + ghdp::b9 * A[9].z + ghdp::b10 * A[10].z + ghdp::b11 * A[11].z + ghdp::b12 * A[12].z;
    pn.z = p0.z + h * kp;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(p0.z), fabs(pn.z)));
    double sqr = (kp - ghdp::bhh1 * A[1].z - ghdp::bhh2 * A[9].z - ghdp::bhh3 * A[12].z) * isk;
    err3 += sqr * sqr;
    sqr = (ghdp::er1 * A[1].z + ghdp::er6 * A[6].z + ghdp::er7 * A[7].z + ghdp::er8 * A[8].z + ghdp::er9 * A[9].z + ghdp::er10 * A[10].z + ghdp::er11 * A[11].z + ghdp::er12 * A[12].z) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kp = ghdp::b1 * A[1].dx + ghdp::b6 * A[6].dx + ghdp::b7 * A[7].dx + ghdp::b8 * A[8].dx + ghdp::b9 * A[9].dx + ghdp::b10 * A[10].dx + ghdp::b11 * A[11].dx + ghdp::b12 * A[12].dx;
    pn.dx = p0.dx + h * kp;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(p0.dx), fabs(pn.dx)));
    double sqr = (kp - ghdp::bhh1 * A[1].dx - ghdp::bhh2 * A[9].dx - ghdp::bhh3 * A[12].dx) * isk;
    err3 += sqr * sqr;
    sqr = (ghdp::er1 * A[1].dx + ghdp::er6 * A[6].dx + ghdp::er7 * A[7].dx + ghdp::er8 * A[8].dx + ghdp::er9 * A[9].dx + ghdp::er10 * A[10].dx + ghdp::er11 * A[11].dx + ghdp::er12 * A[12].dx) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kp = ghdp::b1 * A[1].dy + ghdp::b6 * A[6].dy + ghdp::b7 * A[7].dy + ghdp::b8 * A[8].dy + ghdp::b9 * A[9].dy + ghdp::b10 * A[10].dy + ghdp::b11 * A[11].dy + ghdp::b12 * A[12].dy;
    pn.dy = p0.dy + h * kp;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(p0.dy), fabs(pn.dy)));
    double sqr = (kp - ghdp::bhh1 * A[1].dy - ghdp::bhh2 * A[9].dy - ghdp::bhh3 * A[12].dy) * isk;
    err3 += sqr * sqr;
    sqr = (ghdp::er1 * A[1].dy + ghdp::er6 * A[6].dy + ghdp::er7 * A[7].dy + ghdp::er8 * A[8].dy + ghdp::er9 * A[9].dy + ghdp::er10 * A[10].dy + ghdp::er11 * A[11].dy + ghdp::er12 * A[12].dy) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kp = ghdp::b1 * A[1].dz + ghdp::b6 * A[6].dz + ghdp::b7 * A[7].dz + ghdp::b8 * A[8].dz + ghdp::b9 * A[9].dz + ghdp::b10 * A[10].dz + ghdp::b11 * A[11].dz + ghdp::b12 * A[12].dz;
    pn.dz = p0.dz + h * kp;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(p0.dz), fabs(pn.dz)));
    double sqr = (kp - ghdp::bhh1 * A[1].dz - ghdp::bhh2 * A[9].dz - ghdp::bhh3 * A[12].dz) * isk;
    err3 += sqr * sqr;
    sqr = (ghdp::er1 * A[1].dz + ghdp::er6 * A[6].dz + ghdp::er7 * A[7].dz + ghdp::er8 * A[8].dz + ghdp::er9 * A[9].dz + ghdp::er10 * A[10].dz + ghdp::er11 * A[11].dz + ghdp::er12 * A[12].dz) * isk;
    err5 += sqr * sqr;
  }
}

// Dense output (leung_dop853.c:430-474) of the accepted step at the time
// t_old + sd h (s1 = 1 - sd), after the dense stages 14-16: per component
// the coefficients rc1..rc8 of the continuous extension, then
// rc1 + sd (rc2 + s1 (rc3 + sd (rc4 + s1 (rc5 + sd (rc6 + s1 (rc7 + sd rc8)))))).
__device__ __forceinline__ void gh_rkn_dense_eval(double h, double sd, double s1, const GhVec6& q0,
                                                 const GhVec6& p0, const GhVec6& qn,
                                                 const GhVec6& pn, const GhVec6* A, double* yo) {
  {
    const double ydiff = qn.x - q0.x;
    const double bspl = h * p0.x - ydiff;
    const double r0 = q0.x, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * pn.x - bspl;
    const double r4 = h * (ghrk::dp4 * p0.x + ghrk::dn4 * pn.x + h * (ghrk::da4_1 * A[1].x + ghrk::da4_4 * A[4].x + ghrk::da4_5 * A[5].x + ghrk::da4_6 * A[6].x + ghrk::da4_7 * A[7].x + ghrk::da4_8 * A[8].x + ghrk::da4_9 * A[9].x + ghrk::da4_10 * A[10].x + ghrk::da4_11 * A[11].x + ghrk::da4_12 * A[12].x + ghrk::da4_13 * A[13].x + ghrk::da4_14 * A[14].x + ghrk::da4_15 * A[15].x));
    const double r5 = h * (ghrk::dp5 * p0.x + ghrk::dn5 * pn.x + h * (ghrk::da5_1 * A[1].x + ghrk::da5_4 * A[4].x + ghrk::da5_5 * A[5].x + ghrk::da5_6 * A[6].x + ghrk::da5_7 * A[7].x + ghrk::da5_8 * A[8].x + ghrk::da5_9 * A[9].x + ghrk::da5_10 * A[10].x + ghrk::da5_11 * A[11].x + ghrk::da5_12 * A[12].x + ghrk::da5_13 * A[13].x + ghrk::da5_14 * A[14].x + ghrk::da5_15 * A[15].x));
    const double r6 = h * (ghrk::dp6 * p0.x + ghrk::dn6 * pn.x + h * (ghrk::da6_1 * A[1].x + ghrk::da6_4 * A[4].x + ghrk::da6_5 * A[5].x + ghrk::da6_6 * A[6].x + ghrk::da6_7 * A[7].x + ghrk::da6_8 * A[8].x + ghrk::da6_9 * A[9].x + ghrk::da6_10 * A[10].x + ghrk::da6_11 * A[11].x + ghrk::da6_12 * A[12].x + ghrk::da6_13 * A[13].x + ghrk::da6_14 * A[14].x + ghrk::da6_15 * A[15].x));
    const double r7 = h * (ghrk::dp7 * p0.x + ghrk::dn7 * pn.x + h * (ghrk::da7_1 * A[1].x + ghrk::da7_4 * A[4].x + ghrk::da7_5 * A[5].x + ghrk::da7_6 * A[6].x + ghrk::da7_7 * A[7].x + ghrk::da7_8 * A[8].x + ghrk::da7_9 * A[9].x + ghrk::da7_10 * A[10].x + ghrk::da7_11 * A[11].x + ghrk::da7_12 * A[12].x + ghrk::da7_13 * A[13].x + ghrk::da7_14 * A[14].x + ghrk::da7_15 * A[15].x));
    yo[0] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = qn.y - q0.y;
    const double bspl = h * p0.y - ydiff;
    const double r0 = q0.y, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * pn.y - bspl;
    const double r4 = h * (ghrk::dp4 * p0.y + ghrk::dn4 * pn.y + h * (ghrk::da4_1 * A[1].y + ghrk::da4_4 * A[4].y + ghrk::da4_5 * A[5].y + ghrk::da4_6 * A[6].y + ghrk::da4_7 * A[7].y + ghrk::da4_8 * A[8].y + ghrk::da4_9 * A[9].y + ghrk::da4_10 * A[10].y + ghrk::da4_11 * A[11].y + ghrk::da4_12 * A[12].y + ghrk::da4_13 * A[13].y + ghrk::da4_14 * A[14].y + ghrk::da4_15 * A[15].y));
    const double r5 = h * (ghrk::dp5 * p0.y + ghrk::dn5 * pn.y + h * (ghrk::da5_1 * A[1].y + ghrk::da5_4 * A[4].y + ghrk::da5_5 * A[5].y + ghrk::da5_6 * A[6].y + ghrk::da5_7 * A[7].y + ghrk::da5_8 * A[8].y + ghrk::da5_9 * A[9].y + ghrk::da5_10 * A[10].y + ghrk::da5_11 * A[11].y + ghrk::da5_12 * A[12].y + ghrk::da5_13 * A[13].y + ghrk::da5_14 * A[14].y + ghrk::da5_15 * A[15].y));
    const double r6 = h * (ghrk::dp6 * p0.y + ghrk::dn6 * pn.y + h * (ghrk::da6_1 * A[1].y + ghrk::da6_4 * A[4].y + ghrk::da6_5 * A[5].y + ghrk::da6_6 * A[6].y + ghrk::da6_7 * A[7].y + ghrk::da6_8 * A[8].y + ghrk::da6_9 * A[9].y + ghrk::da6_10 * A[10].y + ghrk::da6_11 * A[11].y + ghrk::da6_12 * A[12].y + ghrk::da6_13 * A[13].y + ghrk::da6_14 * A[14].y + ghrk::da6_15 * A[15].y));
    const double r7 = h * (ghrk::dp7 * p0.y + ghrk::dn7 * pn.y + h * (ghrk::da7_1 * A[1].y + ghrk::da7_4 * A[4].y + ghrk::da7_5 * A[5].y + ghrk::da7_6 * A[6].y + ghrk::da7_7 * A[7].y + ghrk::da7_8 * A[8].y + ghrk::da7_9 * A[9].y + ghrk::da7_10 * A[10].y + ghrk::da7_11 * A[11].y + ghrk::da7_12 * A[12].y + ghrk::da7_13 * A[13].y + ghrk::da7_14 * A[14].y + ghrk::da7_15 * A[15].y));
    yo[1] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = qn.z - q0.z;
    const double bspl = h * p0.z - ydiff;
    const double r0 = q0.z, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * pn.z - bspl;
    const double r4 = h * (ghrk::dp4 * p0.z + ghrk::dn4 * pn.z + h * (ghrk::da4_1 * A[1].z + ghrk::da4_4 * A[4].z + ghrk::da4_5 * A[5].z + ghrk::da4_6 * A[6].z + ghrk::da4_7 * A[7].z + ghrk::da4_8 * A[8].z + ghrk::da4_9 * A[9].z + ghrk::da4_10 * A[10].z + ghrk::da4_11 * A[11].z + ghrk::da4_12 * A[12].z + ghrk::da4_13 * A[13].z + ghrk::da4_14 * A[14].z + ghrk::da4_15 * A[15].z));
    const double r5 = h * (ghrk::dp5 * p0.z + ghrk::dn5 * pn.z + h * (ghrk::da5_1 * A[1].z + ghrk::da5_4 * A[4].z + ghrk::da5_5 * A[5].z + ghrk::da5_6 * A[6].z + ghrk::da5_7 * A[7].z + ghrk::da5_8 * A[8].z + ghrk::da5_9 * A[9].z + ghrk::da5_10 * A[10].z + ghrk::da5_11 * A[11].z + ghrk::da5_12 * A[12].z + ghrk::da5_13 * A[13].z + ghrk::da5_14 * A[14].z + ghrk::da5_15 * A[15].z));
    const double r6 = h * (ghrk::dp6 * p0.z + ghrk::dn6 * pn.z + h * (ghrk::da6_1 * A[1].z + ghrk::da6_4 * A[4].z + ghrk::da6_5 * A[5].z + ghrk::da6_6 * A[6].z + ghrk::da6_7 * A[7].z + ghrk::da6_8 * A[8].z + ghrk::da6_9 * A[9].z + ghrk::da6_10 * A[10].z + ghrk::da6_11 * A[11].z + ghrk::da6_12 * A[12].z + ghrk::da6_13 * A[13].z + ghrk::da6_14 * A[14].z + ghrk::da6_15 * A[15].z));
    const double r7 = h * (ghrk::dp7 * p0.z + ghrk::dn7 * pn.z + h * (ghrk::da7_1 * A[1].z + ghrk::da7_4 * A[4].z + ghrk::da7_5 * A[5].z + ghrk::da7_6 * A[6].z + ghrk::da7_7 * A[7].z + ghrk::da7_8 * A[8].z + ghrk::da7_9 * A[9].z + ghrk::da7_10 * A[10].z + ghrk::da7_11 * A[11].z + ghrk::da7_12 * A[12].z + ghrk::da7_13 * A[13].z + ghrk::da7_14 * A[14].z + ghrk::da7_15 * A[15].z));
    yo[2] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = qn.dx - q0.dx;
    const double bspl = h * p0.dx - ydiff;
    const double r0 = q0.dx, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * pn.dx - bspl;
    const double r4 = h * (ghrk::dp4 * p0.dx + ghrk::dn4 * pn.dx + h * (ghrk::da4_1 * A[1].dx + ghrk::da4_4 * A[4].dx + ghrk::da4_5 * A[5].dx + ghrk::da4_6 * A[6].dx + ghrk::da4_7 * A[7].dx + ghrk::da4_8 * A[8].dx + ghrk::da4_9 * A[9].dx + ghrk::da4_10 * A[10].dx + ghrk::da4_11 * A[11].dx + ghrk::da4_12 * A[12].dx + ghrk::da4_13 * A[13].dx + ghrk::da4_14 * A[14].dx + ghrk::da4_15 * A[15].dx));
    const double r5 = h * (ghrk::dp5 * p0.dx + ghrk::dn5 * pn.dx + h * (ghrk::da5_1 * A[1].dx + ghrk::da5_4 * A[4].dx + ghrk::da5_5 * A[5].dx + ghrk::da5_6 * A[6].dx + ghrk::da5_7 * A[7].dx + ghrk::da5_8 * A[8].dx + ghrk::da5_9 * A[9].dx + ghrk::da5_10 * A[10].dx + ghrk::da5_11 * A[11].dx + ghrk::da5_12 * A[12].dx + ghrk::da5_13 * A[13].dx + ghrk::da5_14 * A[14].dx + ghrk::da5_15 * A[15].dx));
    const double r6 = h * (ghrk::dp6 * p0.dx + ghrk::dn6 * pn.dx + h * (ghrk::da6_1 * A[1].dx + ghrk::da6_4 * A[4].dx + ghrk::da6_5 * A[5].dx + ghrk::da6_6 * A[6].dx + ghrk::da6_7 * A[7].dx + ghrk::da6_8 * A[8].dx + ghrk::da6_9 * A[9].dx + ghrk::da6_10 * A[10].dx + ghrk::da6_11 * A[11].dx + ghrk::da6_12 * A[12].dx + ghrk::da6_13 * A[13].dx + ghrk::da6_14 * A[14].dx + ghrk::da6_15 * A[15].dx));
    const double r7 = h * (ghrk::dp7 * p0.dx + ghrk::dn7 * pn.dx + h * (ghrk::da7_1 * A[1].dx + ghrk::da7_4 * A[4].dx + ghrk::da7_5 * A[5].dx + ghrk::da7_6 * A[6].dx + ghrk::da7_7 * A[7].dx + ghrk::da7_8 * A[8].dx + ghrk::da7_9 * A[9].dx + ghrk::da7_10 * A[10].dx + ghrk::da7_11 * A[11].dx + ghrk::da7_12 * A[12].dx + ghrk::da7_13 * A[13].dx + ghrk::da7_14 * A[14].dx + ghrk::da7_15 * A[15].dx));
    yo[3] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = qn.dy - q0.dy;
    const double bspl = h * p0.dy - ydiff;
    const double r0 = q0.dy, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * pn.dy - bspl;
    const double r4 = h * (ghrk::dp4 * p0.dy + ghrk::dn4 * pn.dy + h * (ghrk::da4_1 * A[1].dy + ghrk::da4_4 * A[4].dy + ghrk::da4_5 * A[5].dy + ghrk::da4_6 * A[6].dy + ghrk::da4_7 * A[7].dy + ghrk::da4_8 * A[8].dy + ghrk::da4_9 * A[9].dy + ghrk::da4_10 * A[10].dy + ghrk::da4_11 * A[11].dy + ghrk::da4_12 * A[12].dy + ghrk::da4_13 * A[13].dy + ghrk::da4_14 * A[14].dy + ghrk::da4_15 * A[15].dy));
    const double r5 = h * (ghrk::dp5 * p0.dy + ghrk::dn5 * pn.dy + h * (ghrk::da5_1 * A[1].dy + ghrk::da5_4 * A[4].dy + ghrk::da5_5 * A[5].dy + ghrk::da5_6 * A[6].dy + ghrk::da5_7 * A[7].dy + ghrk::da5_8 * A[8].dy + ghrk::da5_9 * A[9].dy + ghrk::da5_10 * A[10].dy + ghrk::da5_11 * A[11].dy + ghrk::da5_12 * A[12].dy + ghrk::da5_13 * A[13].dy + ghrk::da5_14 * A[14].dy + ghrk::da5_15 * A[15].dy));
    const double r6 = h * (ghrk::dp6 * p0.dy + ghrk::dn6 * pn.dy + h * (ghrk::da6_1 * A[1].dy + ghrk::da6_4 * A[4].dy + ghrk::da6_5 * A[5].dy + ghrk::da6_6 * A[6].dy + ghrk::da6_7 * A[7].dy + ghrk::da6_8 * A[8].dy + ghrk::da6_9 * A[9].dy + ghrk::da6_10 * A[10].dy + ghrk::da6_11 * A[11].dy + ghrk::da6_12 * A[12].dy + ghrk::da6_13 * A[13].dy + ghrk::da6_14 * A[14].dy + ghrk::da6_15 * A[15].dy));
    const double r7 = h * (ghrk::dp7 * p0.dy + ghrk::dn7 * pn.dy + h * (ghrk::da7_1 * A[1].dy + ghrk::da7_4 * A[4].dy + ghrk::da7_5 * A[5].dy + ghrk::da7_6 * A[6].dy + ghrk::da7_7 * A[7].dy + ghrk::da7_8 * A[8].dy + ghrk::da7_9 * A[9].dy + ghrk::da7_10 * A[10].dy + ghrk::da7_11 * A[11].dy + ghrk::da7_12 * A[12].dy + ghrk::da7_13 * A[13].dy + ghrk::da7_14 * A[14].dy + ghrk::da7_15 * A[15].dy));
    yo[4] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = qn.dz - q0.dz;
    const double bspl = h * p0.dz - ydiff;
    const double r0 = q0.dz, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * pn.dz - bspl;
    const double r4 = h * (ghrk::dp4 * p0.dz + ghrk::dn4 * pn.dz + h * (ghrk::da4_1 * A[1].dz + ghrk::da4_4 * A[4].dz + ghrk::da4_5 * A[5].dz + ghrk::da4_6 * A[6].dz + ghrk::da4_7 * A[7].dz + ghrk::da4_8 * A[8].dz + ghrk::da4_9 * A[9].dz + ghrk::da4_10 * A[10].dz + ghrk::da4_11 * A[11].dz + ghrk::da4_12 * A[12].dz + ghrk::da4_13 * A[13].dz + ghrk::da4_14 * A[14].dz + ghrk::da4_15 * A[15].dz));
    const double r5 = h * (ghrk::dp5 * p0.dz + ghrk::dn5 * pn.dz + h * (ghrk::da5_1 * A[1].dz + ghrk::da5_4 * A[4].dz + ghrk::da5_5 * A[5].dz + ghrk::da5_6 * A[6].dz + ghrk::da5_7 * A[7].dz + ghrk::da5_8 * A[8].dz + ghrk::da5_9 * A[9].dz + ghrk::da5_10 * A[10].dz + ghrk::da5_11 * A[11].dz + ghrk::da5_12 * A[12].dz + ghrk::da5_13 * A[13].dz + ghrk::da5_14 * A[14].dz + ghrk::da5_15 * A[15].dz));
    const double r6 = h * (ghrk::dp6 * p0.dz + ghrk::dn6 * pn.dz + h * (ghrk::da6_1 * A[1].dz + ghrk::da6_4 * A[4].dz + ghrk::da6_5 * A[5].dz + ghrk::da6_6 * A[6].dz + ghrk::da6_7 * A[7].dz + ghrk::da6_8 * A[8].dz + ghrk::da6_9 * A[9].dz + ghrk::da6_10 * A[10].dz + ghrk::da6_11 * A[11].dz + ghrk::da6_12 * A[12].dz + ghrk::da6_13 * A[13].dz + ghrk::da6_14 * A[14].dz + ghrk::da6_15 * A[15].dz));
    const double r7 = h * (ghrk::dp7 * p0.dz + ghrk::dn7 * pn.dz + h * (ghrk::da7_1 * A[1].dz + ghrk::da7_4 * A[4].dz + ghrk::da7_5 * A[5].dz + ghrk::da7_6 * A[6].dz + ghrk::da7_7 * A[7].dz + ghrk::da7_8 * A[8].dz + ghrk::da7_9 * A[9].dz + ghrk::da7_10 * A[10].dz + ghrk::da7_11 * A[11].dz + ghrk::da7_12 * A[12].dz + ghrk::da7_13 * A[13].dz + ghrk::da7_14 * A[14].dz + ghrk::da7_15 * A[15].dz));
    yo[5] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = pn.x - p0.x;
    const double bspl = h * A[1].x - ydiff;
    const double r0 = p0.x, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * A[13].x - bspl;
    const double r4 = h * (ghdp::d41 * A[1].x + ghdp::d46 * A[6].x + ghdp::d47 * A[7].x + ghdp::d48 * A[8].x + ghdp::d49 * A[9].x + ghdp::d410 * A[10].x + ghdp::d411 * A[11].x + ghdp::d412 * A[12].x + ghdp::d413 * A[13].x + ghdp::d414 * A[14].x + ghdp::d415 * A[15].x + ghdp::d416 * A[16].x);
    const double r5 = h * (ghdp::d51 * A[1].x + ghdp::d56 * A[6].x + ghdp::d57 * A[7].x + ghdp::d58 * A[8].x + ghdp::d59 * A[9].x + ghdp::d510 * A[10].x + ghdp::d511 * A[11].x + ghdp::d512 * A[12].x + ghdp::d513 * A[13].x + ghdp::d514 * A[14].x + ghdp::d515 * A[15].x + ghdp::d516 * A[16].x);
    const double r6 = h * (ghdp::d61 * A[1].x + ghdp::d66 * A[6].x + ghdp::d67 * A[7].x + ghdp::d68 * A[8].x + ghdp::d69 * A[9].x + ghdp::d610 * A[10].x + ghdp::d611 * A[11].x + ghdp::d612 * A[12].x + ghdp::d613 * A[13].x + ghdp::d614 * A[14].x + ghdp::d615 * A[15].x + ghdp::d616 * A[16].x);
    const double r7 = h * (ghdp::d71 * A[1].x + ghdp::d76 * A[6].x + ghdp::d77 * A[7].x + ghdp::d78 * A[8].x + ghdp::d79 * A[9].x + ghdp::d710 * A[10].x + ghdp::d711 * A[11].x + ghdp::d712 * A[12].x + ghdp::d713 * A[13].x + ghdp::d714 * A[14].x + ghdp::d715 * A[15].x + ghdp::d716 * A[16].x);
    yo[6] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = pn.y - p0.y;
    const double bspl = h * A[1].y - ydiff;
    const double r0 = p0.y, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * A[13].y - bspl;
    const double r4 = h * (ghdp::d41 * A[1].y + ghdp::d46 * A[6].y + ghdp::d47 * A[7].y + ghdp::d48 * A[8].y + ghdp::d49 * A[9].y + ghdp::d410 * A[10].y + ghdp::d411 * A[11].y + ghdp::d412 * A[12].y + ghdp::d413 * A[13].y + ghdp::d414 * A[14].y + ghdp::d415 * A[15].y + ghdp::d416 * A[16].y);
    const double r5 = h * (ghdp::d51 * A[1].y + ghdp::d56 * A[6].y + ghdp::d57 * A[7].y + ghdp::d58 * A[8].y + ghdp::d59 * A[9].y + ghdp::d510 * A[10].y + ghdp::d511 * A[11].y + ghdp::d512 * A[12].y + ghdp::d513 * A[13].y + ghdp::d514 * A[14].y + ghdp::d515 * A[15].y + ghdp::d516 * A[16].y);
    const double r6 = h * (ghdp::d61 * A[1].y + ghdp::d66 * A[6].y + ghdp::d67 * A[7].y + ghdp::d68 * A[8].y + ghdp::d69 * A[9].y + ghdp::d610 * A[10].y + ghdp::d611 * A[11].y + ghdp::d612 * A[12].y + ghdp::d613 * A[13].y + ghdp::d614 * A[14].y + ghdp::d615 * A[15].y + ghdp::d616 * A[16].y);
    const double r7 = h * (ghdp::d71 * A[1].y + ghdp::d76 * A[6].y + ghdp::d77 * A[7].y + ghdp::d78 * A[8].y + ghdp::d79 * A[9].y + ghdp::d710 * A[10].y + ghdp::d711 * A[11].y + ghdp::d712 * A[12].y + ghdp::d713 * A[13].y + ghdp::d714 * A[14].y + ghdp::d715 * A[15].y + ghdp::d716 * A[16].y);
    yo[7] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = pn.z - p0.z;
    const double bspl = h * A[1].z - ydiff;
    const double r0 = p0.z, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * A[13].z - bspl;
    const double r4 = h * (ghdp::d41 * A[1].z + ghdp::d46 * A[6].z + ghdp::d47 * A[7].z + ghdp::d48 * A[8].z + ghdp::d49 * A[9].z + ghdp::d410 * A[10].z + ghdp::d411 * A[11].z + ghdp::d412 * A[12].z + ghdp::d413 * A[13].z + ghdp::d414 * A[14].z + ghdp::d415 * A[15].z + ghdp::d416 * A[16].z);
    const double r5 = h * (ghdp::d51 * A[1].z + ghdp::d56 * A[6].z + ghdp::d57 * A[7].z + ghdp::d58 * A[8].z + ghdp::d59 * A[9].z + ghdp::d510 * A[10].z + ghdp::d511 * A[11].z + ghdp::d512 * A[12].z + ghdp::d513 * A[13].z + ghdp::d514 * A[14].z + ghdp::d515 * A[15].z + ghdp::d516 * A[16].z);
    const double r6 = h * (ghdp::d61 * A[1].z + ghdp::d66 * A[6].z + ghdp::d67 * A[7].z + ghdp::d68 * A[8].z + ghdp::d69 * A[9].z + ghdp::d610 * A[10].z + ghdp::d611 * A[11].z + ghdp::d612 * A[12].z + ghdp::d613 * A[13].z + ghdp::d614 * A[14].z + ghdp::d615 * A[15].z + ghdp::d616 * A[16].z);
    const double r7 = h * (ghdp::d71 * A[1].z + ghdp::d76 * A[6].z + ghdp::d77 * A[7].z + ghdp::d78 * A[8].z + ghdp::d79 * A[9].z + ghdp::d710 * A[10].z + ghdp::d711 * A[11].z + ghdp::d712 * A[12].z + ghdp::d713 * A[13].z + ghdp::d714 * A[14].z + ghdp::d715 * A[15].z + ghdp::d716 * A[16].z);
    yo[8] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = pn.dx - p0.dx;
    const double bspl = h * A[1].dx - ydiff;
    const double r0 = p0.dx, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * A[13].dx - bspl;
    const double r4 = h * (ghdp::d41 * A[1].dx + ghdp::d46 * A[6].dx + ghdp::d47 * A[7].dx + ghdp::d48 * A[8].dx + ghdp::d49 * A[9].dx + ghdp::d410 * A[10].dx + ghdp::d411 * A[11].dx + ghdp::d412 * A[12].dx + ghdp::d413 * A[13].dx + ghdp::d414 * A[14].dx + ghdp::d415 * A[15].dx + ghdp::d416 * A[16].dx);
    const double r5 = h * (ghdp::d51 * A[1].dx + ghdp::d56 * A[6].dx + ghdp::d57 * A[7].dx + ghdp::d58 * A[8].dx + ghdp::d59 * A[9].dx + ghdp::d510 * A[10].dx + ghdp::d511 * A[11].dx + ghdp::d512 * A[12].dx + ghdp::d513 * A[13].dx + ghdp::d514 * A[14].dx + ghdp::d515 * A[15].dx + ghdp::d516 * A[16].dx);
    const double r6 = h * (ghdp::d61 * A[1].dx + ghdp::d66 * A[6].dx + ghdp::d67 * A[7].dx + ghdp::d68 * A[8].dx + ghdp::d69 * A[9].dx + ghdp::d610 * A[10].dx + ghdp::d611 * A[11].dx + ghdp::d612 * A[12].dx + ghdp::d613 * A[13].dx + ghdp::d614 * A[14].dx + ghdp::d615 * A[15].dx + ghdp::d616 * A[16].dx);
    const double r7 = h * (ghdp::d71 * A[1].dx + ghdp::d76 * A[6].dx + ghdp::d77 * A[7].dx + ghdp::d78 * A[8].dx + ghdp::d79 * A[9].dx + ghdp::d710 * A[10].dx + ghdp::d711 * A[11].dx + ghdp::d712 * A[12].dx + ghdp::d713 * A[13].dx + ghdp::d714 * A[14].dx + ghdp::d715 * A[15].dx + ghdp::d716 * A[16].dx);
    yo[9] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = pn.dy - p0.dy;
    const double bspl = h * A[1].dy - ydiff;
    const double r0 = p0.dy, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * A[13].dy - bspl;
    const double r4 = h * (ghdp::d41 * A[1].dy + ghdp::d46 * A[6].dy + ghdp::d47 * A[7].dy + ghdp::d48 * A[8].dy + ghdp::d49 * A[9].dy + ghdp::d410 * A[10].dy + ghdp::d411 * A[11].dy + ghdp::d412 * A[12].dy + ghdp::d413 * A[13].dy + ghdp::d414 * A[14].dy + ghdp::d415 * A[15].dy + ghdp::d416 * A[16].dy);
    const double r5 = h * (ghdp::d51 * A[1].dy + ghdp::d56 * A[6].dy + ghdp::d57 * A[7].dy + ghdp::d58 * A[8].dy + ghdp::d59 * A[9].dy + ghdp::d510 * A[10].dy + ghdp::d511 * A[11].dy + ghdp::d512 * A[12].dy + ghdp::d513 * A[13].dy + ghdp::d514 * A[14].dy + ghdp::d515 * A[15].dy + ghdp::d516 * A[16].dy);
    const double r6 = h * (ghdp::d61 * A[1].dy + ghdp::d66 * A[6].dy + ghdp::d67 * A[7].dy + ghdp::d68 * A[8].dy + ghdp::d69 * A[9].dy + ghdp::d610 * A[10].dy + ghdp::d611 * A[11].dy + ghdp::d612 * A[12].dy + ghdp::d613 * A[13].dy + ghdp::d614 * A[14].dy + ghdp::d615 * A[15].dy + ghdp::d616 * A[16].dy);
    const double r7 = h * (ghdp::d71 * A[1].dy + ghdp::d76 * A[6].dy + ghdp::d77 * A[7].dy + ghdp::d78 * A[8].dy + ghdp::d79 * A[9].dy + ghdp::d710 * A[10].dy + ghdp::d711 * A[11].dy + ghdp::d712 * A[12].dy + ghdp::d713 * A[13].dy + ghdp::d714 * A[14].dy + ghdp::d715 * A[15].dy + ghdp::d716 * A[16].dy);
    yo[10] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = pn.dz - p0.dz;
    const double bspl = h * A[1].dz - ydiff;
    const double r0 = p0.dz, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * A[13].dz - bspl;
    const double r4 = h * (ghdp::d41 * A[1].dz + ghdp::d46 * A[6].dz + ghdp::d47 * A[7].dz + ghdp::d48 * A[8].dz + ghdp::d49 * A[9].dz + ghdp::d410 * A[10].dz + ghdp::d411 * A[11].dz + ghdp::d412 * A[12].dz + ghdp::d413 * A[13].dz + ghdp::d414 * A[14].dz + ghdp::d415 * A[15].dz + ghdp::d416 * A[16].dz);
    const double r5 = h * (ghdp::d51 * A[1].dz + ghdp::d56 * A[6].dz + ghdp::d57 * A[7].dz + ghdp::d58 * A[8].dz + ghdp::d59 * A[9].dz + ghdp::d510 * A[10].dz + ghdp::d511 * A[11].dz + ghdp::d512 * A[12].dz + ghdp::d513 * A[13].dz + ghdp::d514 * A[14].dz + ghdp::d515 * A[15].dz + ghdp::d516 * A[16].dz);
    const double r6 = h * (ghdp::d61 * A[1].dz + ghdp::d66 * A[6].dz + ghdp::d67 * A[7].dz + ghdp::d68 * A[8].dz + ghdp::d69 * A[9].dz + ghdp::d610 * A[10].dz + ghdp::d611 * A[11].dz + ghdp::d612 * A[12].dz + ghdp::d613 * A[13].dz + ghdp::d614 * A[14].dz + ghdp::d615 * A[15].dz + ghdp::d616 * A[16].dz);
    const double r7 = h * (ghdp::d71 * A[1].dz + ghdp::d76 * A[6].dz + ghdp::d77 * A[7].dz + ghdp::d78 * A[8].dz + ghdp::d79 * A[9].dz + ghdp::d710 * A[10].dz + ghdp::d711 * A[11].dz + ghdp::d712 * A[12].dz + ghdp::d713 * A[13].dz + ghdp::d714 * A[14].dz + ghdp::d715 * A[15].dz + ghdp::d716 * A[16].dz);
    yo[11] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
}
// position argument of stage 2
__device__ __forceinline__ GhVec4 gh_rkn_q2(double h, const GhVec4& q0, const GhVec4& p0,
                                             const GhVec4& pn, const GhVec4* A) {
  (void)pn;
  GhVec4 Q;
  Q.x = q0.x + h * (ghrk::qp2 * p0.x);
  Q.y = q0.y + h * (ghrk::qp2 * p0.y);
  Q.dx = q0.dx + h * (ghrk::qp2 * p0.dx);
  Q.dy = q0.dy + h * (ghrk::qp2 * p0.dy);
  return Q;
}

// position argument of stage 3
__device__ __forceinline__ GhVec4 gh_rkn_q3(double h, const GhVec4& q0, const GhVec4& p0,
                                             const GhVec4& pn, const GhVec4* A) {
  (void)pn;
  GhVec4 Q;
  Q.x = q0.x + h * (ghrk::qp3 * p0.x + h * (ghrk::qa3_1 * A[1].x));
  Q.y = q0.y + h * (ghrk::qp3 * p0.y + h * (ghrk::qa3_1 * A[1].y));
  Q.dx = q0.dx + h * (ghrk::qp3 * p0.dx + h * (ghrk::qa3_1 * A[1].dx));
  Q.dy = q0.dy + h * (ghrk::qp3 * p0.dy + h * (ghrk::qa3_1 * A[1].dy));
  return Q;
}

// position argument of stage 4
__device__ __forceinline__ GhVec4 gh_rkn_q4(double h, const GhVec4& q0, const GhVec4& p0,
                                             const GhVec4& pn, const GhVec4* A) {
  (void)pn;
  GhVec4 Q;
  Q.x = q0.x + h * (ghrk::qp4 * p0.x + h * (ghrk::qa4_1 * A[1].x + ghrk::qa4_2 * A[2].x));
  Q.y = q0.y + h * (ghrk::qp4 * p0.y + h * (ghrk::qa4_1 * A[1].y + ghrk::qa4_2 * A[2].y));
  Q.dx = q0.dx + h * (ghrk::qp4 * p0.dx + h * (ghrk::qa4_1 * A[1].dx + ghrk::qa4_2 * A[2].dx));
  Q.dy = q0.dy + h * (ghrk::qp4 * p0.dy + h * (ghrk::qa4_1 * A[1].dy + ghrk::qa4_2 * A[2].dy));
  return Q;
}

// position argument of stage 5
__device__ __forceinline__ GhVec4 gh_rkn_q5(double h, const GhVec4& q0, const GhVec4& p0,
                                             const GhVec4& pn, const GhVec4* A) {
  (void)pn;
  GhVec4 Q;
  Q.x = q0.x + h * (ghrk::qp5 * p0.x + h * (ghrk::qa5_1 * A[1].x + ghrk::qa5_2 * A[2].x + ghrk::qa5_3 * A[3].x));
  Q.y = q0.y + h * (ghrk::qp5 * p0.y + h * (ghrk::qa5_1 * A[1].y + ghrk::qa5_2 * A[2].y + ghrk::qa5_3 * A[3].y));
  Q.dx = q0.dx + h * (ghrk::qp5 * p0.dx + h * (ghrk::qa5_1 * A[1].dx + ghrk::qa5_2 * A[2].dx + ghrk::qa5_3 * A[3].dx));
  Q.dy = q0.dy + h * (ghrk::qp5 * p0.dy + h * (ghrk::qa5_1 * A[1].dy + ghrk::qa5_2 * A[2].dy + ghrk::qa5_3 * A[3].dy));
  return Q;
}

// position argument of stage 6
__device__ __forceinline__ GhVec4 gh_rkn_q6(double h, const GhVec4& q0, const GhVec4& p0,
                                             const GhVec4& pn, const GhVec4* A) {
  (void)pn;
  GhVec4 Q;
  Q.x = q0.x + h * (ghrk::qp6 * p0.x + h * (ghrk::qa6_1 * A[1].x + ghrk::qa6_3 * A[3].x + ghrk::qa6_4 * A[4].x));
  Q.y = q0.y + h * (ghrk::qp6 * p0.y + h * (ghrk::qa6_1 * A[1].y + ghrk::qa6_3 * A[3].y + ghrk::qa6_4 * A[4].y));
  Q.dx = q0.dx + h * (ghrk::qp6 * p0.dx + h * (ghrk::qa6_1 * A[1].dx + ghrk::qa6_3 * A[3].dx + ghrk::qa6_4 * A[4].dx));
  Q.dy = q0.dy + h * (ghrk::qp6 * p0.dy + h * (ghrk::qa6_1 * A[1].dy + ghrk::qa6_3 * A[3].dy + ghrk::qa6_4 * A[4].dy));
  return Q;
}

// position argument of stage 7
__device__ __forceinline__ GhVec4 gh_rkn_q7(double h, const GhVec4& q0, const GhVec4& p0,
                                             const GhVec4& pn, const GhVec4* A) {
  (void)pn;
  GhVec4 Q;
  Q.x = q0.x + h * (ghrk::qp7 * p0.x + h * (ghrk::qa7_1 * A[1].x + ghrk::qa7_3 * A[3].x + ghrk::qa7_4 * A[4].x + ghrk::qa7_5 * A[5].x));
  Q.y = q0.y + h * (ghrk::qp7 * p0.y + h * (ghrk::qa7_1 * A[1].y + ghrk::qa7_3 * A[3].y + ghrk::qa7_4 * A[4].y + ghrk::qa7_5 * A[5].y));
  Q.dx = q0.dx + h * (ghrk::qp7 * p0.dx + h * (ghrk::qa7_1 * A[1].dx + ghrk::qa7_3 * A[3].dx + ghrk::qa7_4 * A[4].dx + ghrk::qa7_5 * A[5].dx));
  Q.dy = q0.dy + h * (ghrk::qp7 * p0.dy + h * (ghrk::qa7_1 * A[1].dy + ghrk::qa7_3 * A[3].dy + ghrk::qa7_4 * A[4].dy + ghrk::qa7_5 * A[5].dy));
  return Q;
}

// position argument of stage 8
__device__ __forceinline__ GhVec4 gh_rkn_q8(double h, const GhVec4& q0, const GhVec4& p0,
                                             const GhVec4& pn, const GhVec4* A) {
  (void)pn;
  GhVec4 Q;
  Q.x = q0.x + h * (ghrk::qp8 * p0.x + h * (ghrk::qa8_1 * A[1].x + ghrk::qa8_3 * A[3].x + ghrk::qa8_4 * A[4].x + ghrk::qa8_5 * A[5].x + ghrk::qa8_6 * A[6].x));
  Q.y = q0.y + h * (ghrk::qp8 * p0.y + h * (ghrk::qa8_1 * A[1].y + ghrk::qa8_3 * A[3].y + ghrk::qa8_4 * A[4].y + ghrk::qa8_5 * A[5].y + ghrk::qa8_6 * A[6].y));
  Q.dx = q0.dx + h * (ghrk::qp8 * p0.dx + h * (ghrk::qa8_1 * A[1].dx + ghrk::qa8_3 * A[3].dx + ghrk::qa8_4 * A[4].dx + ghrk::qa8_5 * A[5].dx + ghrk::qa8_6 * A[6].dx));
  Q.dy = q0.dy + h * (ghrk::qp8 * p0.dy + h * (ghrk::qa8_1 * A[1].dy + ghrk::qa8_3 * A[3].dy + ghrk::qa8_4 * A[4].dy + ghrk::qa8_5 * A[5].dy + ghrk::qa8_6 * A[6].dy));
  return Q;
}

// position argument of stage 9
__device__ __forceinline__ GhVec4 gh_rkn_q9(double h, const GhVec4& q0, const GhVec4& p0,
                                             const GhVec4& pn, const GhVec4* A) {
  (void)pn;
  GhVec4 Q;
  Q.x = q0.x + h * (ghrk::qp9 * p0.x + h * (ghrk::qa9_1 * A[1].x + ghrk::qa9_3 * A[3].x + ghrk::qa9_4 * A[4].x + ghrk::qa9_5 * A[5].x + ghrk::qa9_6 * A[6].x + ghrk::qa9_7 * A[7].x));
  Q.y = q0.y + h * (ghrk::qp9 * p0.y + h * (ghrk::qa9_1 * A[1].y + ghrk::qa9_3 * A[3].y + ghrk::qa9_4 * A[4].y + ghrk::qa9_5 * A[5].y + ghrk::qa9_6 * A[6].y + ghrk::qa9_7 * A[7].y));
  Q.dx = q0.dx + h * (ghrk::qp9 * p0.dx + h * (ghrk::qa9_1 * A[1].dx + ghrk::qa9_3 * A[3].dx + ghrk::qa9_4 * A[4].dx + ghrk::qa9_5 * A[5].dx + ghrk::qa9_6 * A[6].dx + ghrk::qa9_7 * A[7].dx));
  Q.dy = q0.dy + h * (ghrk::qp9 * p0.dy + h * (ghrk::qa9_1 * A[1].dy + ghrk::qa9_3 * A[3].dy + ghrk::qa9_4 * A[4].dy + ghrk::qa9_5 * A[5].dy + ghrk::qa9_6 * A[6].dy + ghrk::qa9_7 * A[7].dy));
  return Q;
}

// position argument of stage 10
__device__ __forceinline__ GhVec4 gh_rkn_q10(double h, const GhVec4& q0, const GhVec4& p0,
                                             const GhVec4& pn, const GhVec4* A) {
  (void)pn;
  GhVec4 Q;
  Q.x = q0.x + h * (ghrk::qp10 * p0.x + h * (ghrk::qa10_1 * A[1].x + ghrk::qa10_3 * A[3].x + ghrk::qa10_4 * A[4].x + ghrk::qa10_5 * A[5].x + ghrk::qa10_6 * A[6].x + ghrk::qa10_7 * A[7].x + ghrk::qa10_8 * A[8].x));
  Q.y = q0.y + h * (ghrk::qp10 * p0.y + h * (ghrk::qa10_1 * A[1].y + ghrk::qa10_3 * A[3].y + ghrk::qa10_4 * A[4].y + ghrk::qa10_5 * A[5].y + ghrk::qa10_6 * A[6].y + ghrk::qa10_7 * A[7].y + ghrk::qa10_8 * A[8].y));
  Q.dx = q0.dx + h * (ghrk::qp10 * p0.dx + h * (ghrk::qa10_1 * A[1].dx + ghrk::qa10_3 * A[3].dx + ghrk::qa10_4 * A[4].dx + ghrk::qa10_5 * A[5].dx + ghrk::qa10_6 * A[6].dx + ghrk::qa10_7 * A[7].dx + ghrk::qa10_8 * A[8].dx));
  Q.dy = q0.dy + h * (ghrk::qp10 * p0.dy + h * (ghrk::qa10_1 * A[1].dy + ghrk::qa10_3 * A[3].dy + ghrk::qa10_4 * A[4].dy + ghrk::qa10_5 * A[5].dy + ghrk::qa10_6 * A[6].dy + ghrk::qa10_7 * A[7].dy + ghrk::qa10_8 * A[8].dy));
  return Q;
}

// position argument of stage 11
__device__ __forceinline__ GhVec4 gh_rkn_q11(double h, const GhVec4& q0, const GhVec4& p0,
                                             const GhVec4& pn, const GhVec4* A) {
  (void)pn;
  GhVec4 Q;
  Q.x = q0.x + h * (ghrk::qp11 * p0.x + h * (ghrk::qa11_1 * A[1].x + ghrk::qa11_3 * A[3].x + ghrk::qa11_4 * A[4].x + ghrk::qa11_5 * A[5].x + ghrk::qa11_6 * A[6].x + ghrk::qa11_7 * A[7].x + ghrk::qa11_8 * A[8].x + ghrk::qa11_9 * A[9].x));
  Q.y = q0.y + h * (ghrk::qp11 * p0.y + h * (ghrk::qa11_1 * A[1].y + ghrk::qa11_3 * A[3].y + ghrk::qa11_4 * A[4].y + ghrk::qa11_5 * A[5].y + ghrk::qa11_6 * A[6].y + ghrk::qa11_7 * A[7].y + ghrk::qa11_8 * A[8].y + ghrk::qa11_9 * A[9].y));
  Q.dx = q0.dx + h * (ghrk::qp11 * p0.dx + h * (ghrk::qa11_1 * A[1].dx + ghrk::qa11_3 * A[3].dx + ghrk::qa11_4 * A[4].dx + ghrk::qa11_5 * A[5].dx + ghrk::qa11_6 * A[6].dx + ghrk::qa11_7 * A[7].dx + ghrk::qa11_8 * A[8].dx + ghrk::qa11_9 * A[9].dx));
  Q.dy = q0.dy + h * (ghrk::qp11 * p0.dy + h * (ghrk::qa11_1 * A[1].dy + ghrk::qa11_3 * A[3].dy + ghrk::qa11_4 * A[4].dy + ghrk::qa11_5 * A[5].dy + ghrk::qa11_6 * A[6].dy + ghrk::qa11_7 * A[7].dy + ghrk::qa11_8 * A[8].dy + ghrk::qa11_9 * A[9].dy));
  return Q;
}

// position argument of stage 12
__device__ __forceinline__ GhVec4 gh_rkn_q12(double h, const GhVec4& q0, const GhVec4& p0,
                                             const GhVec4& pn, const GhVec4* A) {
  (void)pn;
  GhVec4 Q;
  Q.x = q0.x + h * (ghrk::qp12 * p0.x + h * (ghrk::qa12_1 * A[1].x + ghrk::qa12_3 * A[3].x + ghrk::qa12_4 * A[4].x + ghrk::qa12_5 * A[5].x + ghrk::qa12_6 * A[6].x + ghrk::qa12_7 * A[7].x + ghrk::qa12_8 * A[8].x + ghrk::qa12_9 * A[9].x + ghrk::qa12_10 * A[10].x));
  Q.y = q0.y + h * (ghrk::qp12 * p0.y + h * (ghrk::qa12_1 * A[1].y + ghrk::qa12_3 * A[3].y + ghrk::qa12_4 * A[4].y + ghrk::qa12_5 * A[5].y + ghrk::qa12_6 * A[6].y + ghrk::qa12_7 * A[7].y + ghrk::qa12_8 * A[8].y + ghrk::qa12_9 * A[9].y + ghrk::qa12_10 * A[10].y));
  Q.dx = q0.dx + h * (ghrk::qp12 * p0.dx + h * (ghrk::qa12_1 * A[1].dx + ghrk::qa12_3 * A[3].dx + ghrk::qa12_4 * A[4].dx + ghrk::qa12_5 * A[5].dx + ghrk::qa12_6 * A[6].dx + ghrk::qa12_7 * A[7].dx + ghrk::qa12_8 * A[8].dx + ghrk::qa12_9 * A[9].dx + ghrk::qa12_10 * A[10].dx));
  Q.dy = q0.dy + h * (ghrk::qp12 * p0.dy + h * (ghrk::qa12_1 * A[1].dy + ghrk::qa12_3 * A[3].dy + ghrk::qa12_4 * A[4].dy + ghrk::qa12_5 * A[5].dy + ghrk::qa12_6 * A[6].dy + ghrk::qa12_7 * A[7].dy + ghrk::qa12_8 * A[8].dy + ghrk::qa12_9 * A[9].dy + ghrk::qa12_10 * A[10].dy));
  return Q;
}

// position argument of stage 14
__device__ __forceinline__ GhVec4 gh_rkn_q14(double h, const GhVec4& q0, const GhVec4& p0,
                                             const GhVec4& pn, const GhVec4* A) {
  (void)pn;
  GhVec4 Q;
  Q.x = q0.x + h * (ghrk::qp14 * p0.x + ghrk::qn14 * pn.x + h * (ghrk::qa14_1 * A[1].x + ghrk::qa14_4 * A[4].x + ghrk::qa14_5 * A[5].x + ghrk::qa14_6 * A[6].x + ghrk::qa14_7 * A[7].x + ghrk::qa14_8 * A[8].x + ghrk::qa14_9 * A[9].x + ghrk::qa14_10 * A[10].x + ghrk::qa14_11 * A[11].x));
  Q.y = q0.y + h * (ghrk::qp14 * p0.y + ghrk::qn14 * pn.y + h * (ghrk::qa14_1 * A[1].y + ghrk::qa14_4 * A[4].y + ghrk::qa14_5 * A[5].y + ghrk::qa14_6 * A[6].y + ghrk::qa14_7 * A[7].y + ghrk::qa14_8 * A[8].y + ghrk::qa14_9 * A[9].y + ghrk::qa14_10 * A[10].y + ghrk::qa14_11 * A[11].y));
  Q.dx = q0.dx + h * (ghrk::qp14 * p0.dx + ghrk::qn14 * pn.dx + h * (ghrk::qa14_1 * A[1].dx + ghrk::qa14_4 * A[4].dx + ghrk::qa14_5 * A[5].dx + ghrk::qa14_6 * A[6].dx + ghrk::qa14_7 * A[7].dx + ghrk::qa14_8 * A[8].dx + ghrk::qa14_9 * A[9].dx + ghrk::qa14_10 * A[10].dx + ghrk::qa14_11 * A[11].dx));
  Q.dy = q0.dy + h * (ghrk::qp14 * p0.dy + ghrk::qn14 * pn.dy + h * (ghrk::qa14_1 * A[1].dy + ghrk::qa14_4 * A[4].dy + ghrk::qa14_5 * A[5].dy + ghrk::qa14_6 * A[6].dy + ghrk::qa14_7 * A[7].dy + ghrk::qa14_8 * A[8].dy + ghrk::qa14_9 * A[9].dy + ghrk::qa14_10 * A[10].dy + ghrk::qa14_11 * A[11].dy));
  return Q;
}

// position argument of stage 15
__device__ __forceinline__ GhVec4 gh_rkn_q15(double h, const GhVec4& q0, const GhVec4& p0,
                                             const GhVec4& pn, const GhVec4* A) {
  (void)pn;
  GhVec4 Q;
  Q.x = q0.x + h * (ghrk::qp15 * p0.x + ghrk::qn15 * pn.x + h * (ghrk::qa15_1 * A[1].x + ghrk::qa15_4 * A[4].x + ghrk::qa15_5 * A[5].x + ghrk::qa15_6 * A[6].x + ghrk::qa15_7 * A[7].x + ghrk::qa15_8 * A[8].x + ghrk::qa15_9 * A[9].x + ghrk::qa15_10 * A[10].x + ghrk::qa15_11 * A[11].x + ghrk::qa15_12 * A[12].x + ghrk::qa15_13 * A[13].x));
  Q.y = q0.y + h * (ghrk::qp15 * p0.y + ghrk::qn15 * pn.y + h * (ghrk::qa15_1 * A[1].y + ghrk::qa15_4 * A[4].y + ghrk::qa15_5 * A[5].y + ghrk::qa15_6 * A[6].y + ghrk::qa15_7 * A[7].y + ghrk::qa15_8 * A[8].y + ghrk::qa15_9 * A[9].y + ghrk::qa15_10 * A[10].y + ghrk::qa15_11 * A[11].y + ghrk::qa15_12 * A[12].y + ghrk::qa15_13 * A[13].y));
  Q.dx = q0.dx + h * (ghrk::qp15 * p0.dx + ghrk::qn15 * pn.dx + h * (ghrk::qa15_1 * A[1].dx + ghrk::qa15_4 * A[4].dx + ghrk::qa15_5 * A[5].dx + ghrk::qa15_6 * A[6].dx + ghrk::qa15_7 * A[7].dx + ghrk::qa15_8 * A[8].dx + ghrk::qa15_9 * A[9].dx + ghrk::qa15_10 * A[10].dx + ghrk::qa15_11 * A[11].dx + ghrk::qa15_12 * A[12].dx + ghrk::qa15_13 * A[13].dx));
  Q.dy = q0.dy + h * (ghrk::qp15 * p0.dy + ghrk::qn15 * pn.dy + h * (ghrk::qa15_1 * A[1].dy + ghrk::qa15_4 * A[4].dy + ghrk::qa15_5 * A[5].dy + ghrk::qa15_6 * A[6].dy + ghrk::qa15_7 * A[7].dy + ghrk::qa15_8 * A[8].dy + ghrk::qa15_9 * A[9].dy + ghrk::qa15_10 * A[10].dy + ghrk::qa15_11 * A[11].dy + ghrk::qa15_12 * A[12].dy + ghrk::qa15_13 * A[13].dy));
  return Q;
}

// position argument of stage 16
__device__ __forceinline__ GhVec4 gh_rkn_q16(double h, const GhVec4& q0, const GhVec4& p0,
                                             const GhVec4& pn, const GhVec4* A) {
  (void)pn;
  GhVec4 Q;
  Q.x = q0.x + h * (ghrk::qp16 * p0.x + ghrk::qn16 * pn.x + h * (ghrk::qa16_1 * A[1].x + ghrk::qa16_4 * A[4].x + ghrk::qa16_5 * A[5].x + ghrk::qa16_6 * A[6].x + ghrk::qa16_7 * A[7].x + ghrk::qa16_8 * A[8].x + ghrk::qa16_9 * A[9].x + ghrk::qa16_10 * A[10].x + ghrk::qa16_11 * A[11].x + ghrk::qa16_12 * A[12].x + ghrk::qa16_13 * A[13].x + ghrk::qa16_14 * A[14].x));
  Q.y = q0.y + h * (ghrk::qp16 * p0.y + ghrk::qn16 * pn.y + h * (ghrk::qa16_1 * A[1].y + ghrk::qa16_4 * A[4].y + ghrk::qa16_5 * A[5].y + ghrk::qa16_6 * A[6].y + ghrk::qa16_7 * A[7].y + ghrk::qa16_8 * A[8].y + ghrk::qa16_9 * A[9].y + ghrk::qa16_10 * A[10].y + ghrk::qa16_11 * A[11].y + ghrk::qa16_12 * A[12].y + ghrk::qa16_13 * A[13].y + ghrk::qa16_14 * A[14].y));
  Q.dx = q0.dx + h * (ghrk::qp16 * p0.dx + ghrk::qn16 * pn.dx + h * (ghrk::qa16_1 * A[1].dx + ghrk::qa16_4 * A[4].dx + ghrk::qa16_5 * A[5].dx + ghrk::qa16_6 * A[6].dx + ghrk::qa16_7 * A[7].dx + ghrk::qa16_8 * A[8].dx + ghrk::qa16_9 * A[9].dx + ghrk::qa16_10 * A[10].dx + ghrk::qa16_11 * A[11].dx + ghrk::qa16_12 * A[12].dx + ghrk::qa16_13 * A[13].dx + ghrk::qa16_14 * A[14].dx));
  Q.dy = q0.dy + h * (ghrk::qp16 * p0.dy + ghrk::qn16 * pn.dy + h * (ghrk::qa16_1 * A[1].dy + ghrk::qa16_4 * A[4].dy + ghrk::qa16_5 * A[5].dy + ghrk::qa16_6 * A[6].dy + ghrk::qa16_7 * A[7].dy + ghrk::qa16_8 * A[8].dy + ghrk::qa16_9 * A[9].dy + ghrk::qa16_10 * A[10].dy + ghrk::qa16_11 * A[11].dy + ghrk::qa16_12 * A[12].dy + ghrk::qa16_13 * A[13].dy + ghrk::qa16_14 * A[14].dy));
  return Q;
}

// New state (q_new, p_new) and the two error sums of leung_dop853.c:376-404:
// per component, sk = atol + rtol max(|y0|, |y_new|), err3 += (e3/sk)^2,
// err5 += (e5/sk)^2, with e3 = sum (b - bhh) k, e5 = sum er k (one
// reciprocal of sk per component; the norm only steers the step size).
__device__ __forceinline__ void gh_rkn_solution(double h, const GhVec4& q0, const GhVec4& p0,
                                               const GhVec4* A, double atol, double rtol,
                                               GhVec4& qn, GhVec4& pn, double& err5,
                                               double& err3) {
  err5 = 0.0;
  err3 = 0.0;
  {
    const double kq = ghrk::bp * p0.x + h * (ghrk::ba_1 * A[1].x + ghrk::ba_4 * A[4].x + ghrk::ba_5 * A[5].x + ghrk::ba_6 * A[6].x + ghrk::ba_7 * A[7].x + ghrk::ba_8 * A[8].x + ghrk::ba_9 * A[9].x + ghrk::ba_10 * A[10].x + ghrk::ba_11 * A[11].x);
    qn.x = q0.x + h * kq;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(q0.x), fabs(qn.x)));
    double sqr = (ghrk::e3p * p0.x + h * (ghrk::e3a_1 * A[1].x + ghrk::e3a_4 * A[4].x + ghrk::e3a_5 * A[5].x + ghrk::e3a_6 * A[6].x + ghrk::e3a_7 * A[7].x + ghrk::e3a_8 * A[8].x + ghrk::e3a_9 * A[9].x + ghrk::e3a_10 * A[10].x + ghrk::e3a_11 * A[11].x)) * isk;
    err3 += sqr * sqr;
    sqr = (ghrk::e5p * p0.x + h * (ghrk::e5a_1 * A[1].x + ghrk::e5a_4 * A[4].x + ghrk::e5a_5 * A[5].x + ghrk::e5a_6 * A[6].x + ghrk::e5a_7 * A[7].x + ghrk::e5a_8 * A[8].x + ghrk::e5a_9 * A[9].x + ghrk::e5a_10 * A[10].x + ghrk::e5a_11 * A[11].x)) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kq = ghrk::bp * p0.y + h * (ghrk::ba_1 * A[1].y + ghrk::ba_4 * A[4].y + ghrk::ba_5 * A[5].y + ghrk::ba_6 * A[6].y + ghrk::ba_7 * A[7].y + ghrk::ba_8 * A[8].y + ghrk::ba_9 * A[9].y + ghrk::ba_10 * A[10].y + ghrk::ba_11 * A[11].y);
    qn.y = q0.y + h * kq;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(q0.y), fabs(qn.y)));
    double sqr = (ghrk::e3p * p0.y + h * (ghrk::e3a_1 * A[1].y + ghrk::e3a_4 * A[4].y + ghrk::e3a_5 * A[5].y + ghrk::e3a_6 * A[6].y + ghrk::e3a_7 * A[7].y + ghrk::e3a_8 * A[8].y + ghrk::e3a_9 * A[9].y + ghrk::e3a_10 * A[10].y + ghrk::e3a_11 * A[11].y)) * isk;
    err3 += sqr * sqr;
    sqr = (ghrk::e5p * p0.y + h * (ghrk::e5a_1 * A[1].y + ghrk::e5a_4 * A[4].y + ghrk::e5a_5 * A[5].y + ghrk::e5a_6 * A[6].y + ghrk::e5a_7 * A[7].y + ghrk::e5a_8 * A[8].y + ghrk::e5a_9 * A[9].y + ghrk::e5a_10 * A[10].y + ghrk::e5a_11 * A[11].y)) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kq = ghrk::bp * p0.dx + h * (ghrk::ba_1 * A[1].dx + ghrk::ba_4 * A[4].dx + ghrk::ba_5 * A[5].dx + ghrk::ba_6 * A[6].dx + ghrk::ba_7 * A[7].dx + ghrk::ba_8 * A[8].dx + ghrk::ba_9 * A[9].dx + ghrk::ba_10 * A[10].dx + ghrk::ba_11 * A[11].dx);
    qn.dx = q0.dx + h * kq;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(q0.dx), fabs(qn.dx)));
    double sqr = (ghrk::e3p * p0.dx + h * (ghrk::e3a_1 * A[1].dx + ghrk::e3a_4 * A[4].dx + ghrk::e3a_5 * A[5].dx + ghrk::e3a_6 * A[6].dx + ghrk::e3a_7 * A[7].dx + ghrk::e3a_8 * A[8].dx + ghrk::e3a_9 * A[9].dx + ghrk::e3a_10 * A[10].dx + ghrk::e3a_11 * A[11].dx)) * isk;
    err3 += sqr * sqr;
    sqr = (ghrk::e5p * p0.dx + h * (ghrk::e5a_1 * A[1].dx + ghrk::e5a_4 * A[4].dx + ghrk::e5a_5 * A[5].dx + ghrk::e5a_6 * A[6].dx + ghrk::e5a_7 * A[7].dx + ghrk::e5a_8 * A[8].dx + ghrk::e5a_9 * A[9].dx + ghrk::e5a_10 * A[10].dx + ghrk::e5a_11 * A[11].dx)) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kq = ghrk::bp * p0.dy + h * (ghrk::ba_1 * A[1].dy + ghrk::ba_4 * A[4].dy + ghrk::ba_5 * A[5].dy + ghrk::ba_6 * A[6].dy + ghrk::ba_7 * A[7].dy + ghrk::ba_8 * A[8].dy + ghrk::ba_9 * A[9].dy + ghrk::ba_10 * A[10].dy + ghrk::ba_11 * A[11].dy);
    qn.dy = q0.dy + h * kq;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(q0.dy), fabs(qn.dy)));
    double sqr = (ghrk::e3p * p0.dy + h * (ghrk::e3a_1 * A[1].dy + ghrk::e3a_4 * A[4].dy + ghrk::e3a_5 * A[5].dy + ghrk::e3a_6 * A[6].dy + ghrk::e3a_7 * A[7].dy + ghrk::e3a_8 * A[8].dy + ghrk::e3a_9 * A[9].dy + ghrk::e3a_10 * A[10].dy + ghrk::e3a_11 * A[11].dy)) * isk;
    err3 += sqr * sqr;
    sqr = (ghrk::e5p * p0.dy + h * (ghrk::e5a_1 * A[1].dy + ghrk::e5a_4 * A[4].dy + ghrk::e5a_5 * A[5].dy + ghrk::e5a_6 * A[6].dy + ghrk::e5a_7 * A[7].dy + ghrk::e5a_8 * A[8].dy + ghrk::e5a_9 * A[9].dy + ghrk::e5a_10 * A[10].dy + ghrk::e5a_11 * A[11].dy)) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kp = ghdp::b1 * A[1].x + ghdp::b6 * A[6].x + ghdp::b7 * A[7].x + ghdp::b8 * A[8].x + ghdp::b9 * A[9].x + ghdp::b10 * A[10].x + ghdp::b11 * A[11].x + ghdp::b12 * A[12].x;
    pn.x = p0.x + h * kp;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(p0.x), fabs(pn.x)));
    double sqr = (kp - ghdp::bhh1 * A[1].x - ghdp::bhh2 * A[9].x - ghdp::bhh3 * A[12].x) * isk;
    err3 += sqr * sqr;
    sqr = (ghdp::er1 * A[1].x + ghdp::er6 * A[6].x + ghdp::er7 * A[7].x + ghdp::er8 * A[8].x + ghdp::er9 * A[9].x + ghdp::er10 * A[10].x + ghdp::er11 * A[11].x + ghdp::er12 * A[12].x) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kp = ghdp::b1 * A[1].y + ghdp::b6 * A[6].y + ghdp::b7 * A[7].y + ghdp::b8 * A[8].y + ghdp::b9 * A[9].y + ghdp::b10 * A[10].y + ghdp::b11 * A[11].y + ghdp::b12 * A[12].y;
    pn.y = p0.y + h * kp;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(p0.y), fabs(pn.y)));
    double sqr = (kp - ghdp::bhh1 * A[1].y - ghdp::bhh2 * A[9].y - ghdp::bhh3 * A[12].y) * isk;
    err3 += sqr * sqr;
    sqr = (ghdp::er1 * A[1].y + ghdp::er6 * A[6].y + ghdp::er7 * A[7].y + ghdp::er8 * A[8].y + ghdp::er9 * A[9].y + ghdp::er10 * A[10].y + ghdp::er11 * A[11].y + ghdp::er12 * A[12].y) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kp = ghdp::b1 * A[1].dx + ghdp::b6 * A[6].dx + ghdp::b7 * A[7].dx + ghdp::b8 * A[8].dx + ghdp::b9 * A[9].dx + ghdp::b10 * A[10].dx + ghdp::b11 * A[11].dx + ghdp::b12 * A[12].dx;
    pn.dx = p0.dx + h * kp;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(p0.dx), fabs(pn.dx)));
    double sqr = (kp - ghdp::bhh1 * A[1].dx - ghdp::bhh2 * A[9].dx - ghdp::bhh3 * A[12].dx) * isk;
    err3 += sqr * sqr;
    sqr = (ghdp::er1 * A[1].dx + ghdp::er6 * A[6].dx + ghdp::er7 * A[7].dx + ghdp::er8 * A[8].dx + ghdp::er9 * A[9].dx + ghdp::er10 * A[10].dx + ghdp::er11 * A[11].dx + ghdp::er12 * A[12].dx) * isk;
    err5 += sqr * sqr;
  }
  {
    const double kp = ghdp::b1 * A[1].dy + ghdp::b6 * A[6].dy + ghdp::b7 * A[7].dy + ghdp::b8 * A[8].dy + ghdp::b9 * A[9].dy + ghdp::b10 * A[10].dy + ghdp::b11 * A[11].dy + ghdp::b12 * A[12].dy;
    pn.dy = p0.dy + h * kp;
    const double isk = gh_rcp(atol + rtol * fmax(fabs(p0.dy), fabs(pn.dy)));
    double sqr = (kp - ghdp::bhh1 * A[1].dy - ghdp::bhh2 * A[9].dy - ghdp::bhh3 * A[12].dy) * isk;
    err3 += sqr * sqr;
    sqr = (ghdp::er1 * A[1].dy + ghdp::er6 * A[6].dy + ghdp::er7 * A[7].dy + ghdp::er8 * A[8].dy + ghdp::er9 * A[9].dy + ghdp::er10 * A[10].dy + ghdp::er11 * A[11].dy + ghdp::er12 * A[12].dy) * isk;
    err5 += sqr * sqr;
  }
}

// Dense output (leung_dop853.c:430-474) of the accepted step at the time
// t_old + sd h (s1 = 1 - sd), after the dense stages 14-16: per component
// the coefficients rc1..rc8 of the continuous extension, then
// rc1 + sd (rc2 + s1 (rc3 + sd (rc4 + s1 (rc5 + sd (rc6 + s1 (rc7 + sd rc8)))))).
__device__ __forceinline__ void gh_rkn_dense_eval(double h, double sd, double s1, const GhVec4& q0,
                                                 const GhVec4& p0, const GhVec4& qn,
                                                 const GhVec4& pn, const GhVec4* A, double* yo) {
  {
    const double ydiff = qn.x - q0.x;
    const double bspl = h * p0.x - ydiff;
    const double r0 = q0.x, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * pn.x - bspl;
    const double r4 = h * (ghrk::dp4 * p0.x + ghrk::dn4 * pn.x + h * (ghrk::da4_1 * A[1].x + ghrk::da4_4 * A[4].x + ghrk::da4_5 * A[5].x + ghrk::da4_6 * A[6].x + ghrk::da4_7 * A[7].x + ghrk::da4_8 * A[8].x + ghrk::da4_9 * A[9].x + ghrk::da4_10 * A[10].x + ghrk::da4_11 * A[11].x + ghrk::da4_12 * A[12].x + ghrk::da4_13 * A[13].x + ghrk::da4_14 * A[14].x + ghrk::da4_15 * A[15].x));
    const double r5 = h * (ghrk::dp5 * p0.x + ghrk::dn5 * pn.x + h * (ghrk::da5_1 * A[1].x + ghrk::da5_4 * A[4].x + ghrk::da5_5 * A[5].x + ghrk::da5_6 * A[6].x + ghrk::da5_7 * A[7].x + ghrk::da5_8 * A[8].x + ghrk::da5_9 * A[9].x + ghrk::da5_10 * A[10].x + ghrk::da5_11 * A[11].x + ghrk::da5_12 * A[12].x + ghrk::da5_13 * A[13].x + ghrk::da5_14 * A[14].x + ghrk::da5_15 * A[15].x));
    const double r6 = h * (ghrk::dp6 * p0.x + ghrk::dn6 * pn.x + h * (ghrk::da6_1 * A[1].x + ghrk::da6_4 * A[4].x + ghrk::da6_5 * A[5].x + ghrk::da6_6 * A[6].x + ghrk::da6_7 * A[7].x + ghrk::da6_8 * A[8].x + ghrk::da6_9 * A[9].x + ghrk::da6_10 * A[10].x + ghrk::da6_11 * A[11].x + ghrk::da6_12 * A[12].x + ghrk::da6_13 * A[13].x + ghrk::da6_14 * A[14].x + ghrk::da6_15 * A[15].x));
    const double r7 = h * (ghrk::dp7 * p0.x + ghrk::dn7 * pn.x + h * (ghrk::da7_1 * A[1].x + ghrk::da7_4 * A[4].x + ghrk::da7_5 * A[5].x + ghrk::da7_6 * A[6].x + ghrk::da7_7 * A[7].x + ghrk::da7_8 * A[8].x + ghrk::da7_9 * A[9].x + ghrk::da7_10 * A[10].x + ghrk::da7_11 * A[11].x + ghrk::da7_12 * A[12].x + ghrk::da7_13 * A[13].x + ghrk::da7_14 * A[14].x + ghrk::da7_15 * A[15].x));
    yo[0] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = qn.y - q0.y;
    const double bspl = h * p0.y - ydiff;
    const double r0 = q0.y, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * pn.y - bspl;
    const double r4 = h * (ghrk::dp4 * p0.y + ghrk::dn4 * pn.y + h * (ghrk::da4_1 * A[1].y + ghrk::da4_4 * A[4].y + ghrk::da4_5 * A[5].y + ghrk::da4_6 * A[6].y + ghrk::da4_7 * A[7].y + ghrk::da4_8 * A[8].y + ghrk::da4_9 * A[9].y + ghrk::da4_10 * A[10].y + ghrk::da4_11 * A[11].y + ghrk::da4_12 * A[12].y + ghrk::da4_13 * A[13].y + ghrk::da4_14 * A[14].y + ghrk::da4_15 * A[15].y));
    const double r5 = h * (ghrk::dp5 * p0.y + ghrk::dn5 * pn.y + h * (ghrk::da5_1 * A[1].y + ghrk::da5_4 * A[4].y + ghrk::da5_5 * A[5].y + ghrk::da5_6 * A[6].y + ghrk::da5_7 * A[7].y + ghrk::da5_8 * A[8].y + ghrk::da5_9 * A[9].y + ghrk::da5_10 * A[10].y + ghrk::da5_11 * A[11].y + ghrk::da5_12 * A[12].y + ghrk::da5_13 * A[13].y + ghrk::da5_14 * A[14].y + ghrk::da5_15 * A[15].y));
    const double r6 = h * (ghrk::dp6 * p0.y + ghrk::dn6 * pn.y + h * (ghrk::da6_1 * A[1].y + ghrk::da6_4 * A[4].y + ghrk::da6_5 * A[5].y + ghrk::da6_6 * A[6].y + ghrk::da6_7 * A[7].y + ghrk::da6_8 * A[8].y + ghrk::da6_9 * A[9].y + ghrk::da6_10 * A[10].y + ghrk::da6_11 * A[11].y + ghrk::da6_12 * A[12].y + ghrk::da6_13 * A[13].y + ghrk::da6_14 * A[14].y + ghrk::da6_15 * A[15].y));
    const double r7 = h * (ghrk::dp7 * p0.y + ghrk::dn7 * pn.y + h * (ghrk::da7_1 * A[1].y + ghrk::da7_4 * A[4].y + ghrk::da7_5 * A[5].y + ghrk::da7_6 * A[6].y + ghrk::da7_7 * A[7].y + ghrk::da7_8 * A[8].y + ghrk::da7_9 * A[9].y + ghrk::da7_10 * A[10].y + ghrk::da7_11 * A[11].y + ghrk::da7_12 * A[12].y + ghrk::da7_13 * A[13].y + ghrk::da7_14 * A[14].y + ghrk::da7_15 * A[15].y));
    yo[1] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = qn.dx - q0.dx;
    const double bspl = h * p0.dx - ydiff;
    const double r0 = q0.dx, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * pn.dx - bspl;
    const double r4 = h * (ghrk::dp4 * p0.dx + ghrk::dn4 * pn.dx + h * (ghrk::da4_1 * A[1].dx + ghrk::da4_4 * A[4].dx + ghrk::da4_5 * A[5].dx + ghrk::da4_6 * A[6].dx + ghrk::da4_7 * A[7].dx + ghrk::da4_8 * A[8].dx + ghrk::da4_9 * A[9].dx + ghrk::da4_10 * A[10].dx + ghrk::da4_11 * A[11].dx + ghrk::da4_12 * A[12].dx + ghrk::da4_13 * A[13].dx + ghrk::da4_14 * A[14].dx + ghrk::da4_15 * A[15].dx));
    const double r5 = h * (ghrk::dp5 * p0.dx + ghrk::dn5 * pn.dx + h * (ghrk::da5_1 * A[1].dx + ghrk::da5_4 * A[4].dx + ghrk::da5_5 * A[5].dx + ghrk::da5_6 * A[6].dx + ghrk::da5_7 * A[7].dx + ghrk::da5_8 * A[8].dx + ghrk::da5_9 * A[9].dx + ghrk::da5_10 * A[10].dx + ghrk::da5_11 * A[11].dx + ghrk::da5_12 * A[12].dx + ghrk::da5_13 * A[13].dx + ghrk::da5_14 * A[14].dx + ghrk::da5_15 * A[15].dx));
    const double r6 = h * (ghrk::dp6 * p0.dx + ghrk::dn6 * pn.dx + h * (ghrk::da6_1 * A[1].dx + ghrk::da6_4 * A[4].dx + ghrk::da6_5 * A[5].dx + ghrk::da6_6 * A[6].dx + ghrk::da6_7 * A[7].dx + ghrk::da6_8 * A[8].dx + ghrk::da6_9 * A[9].dx + ghrk::da6_10 * A[10].dx + ghrk::da6_11 * A[11].dx + ghrk::da6_12 * A[12].dx + ghrk::da6_13 * A[13].dx + ghrk::da6_14 * A[14].dx + ghrk::da6_15 * A[15].dx));
    const double r7 = h * (ghrk::dp7 * p0.dx + ghrk::dn7 * pn.dx + h * (ghrk::da7_1 * A[1].dx + ghrk::da7_4 * A[4].dx + ghrk::da7_5 * A[5].dx + ghrk::da7_6 * A[6].dx + ghrk::da7_7 * A[7].dx + ghrk::da7_8 * A[8].dx + ghrk::da7_9 * A[9].dx + ghrk::da7_10 * A[10].dx + ghrk::da7_11 * A[11].dx + ghrk::da7_12 * A[12].dx + ghrk::da7_13 * A[13].dx + ghrk::da7_14 * A[14].dx + ghrk::da7_15 * A[15].dx));
    yo[2] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = qn.dy - q0.dy;
    const double bspl = h * p0.dy - ydiff;
    const double r0 = q0.dy, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * pn.dy - bspl;
    const double r4 = h * (ghrk::dp4 * p0.dy + ghrk::dn4 * pn.dy + h * (ghrk::da4_1 * A[1].dy + ghrk::da4_4 * A[4].dy + ghrk::da4_5 * A[5].dy + ghrk::da4_6 * A[6].dy + ghrk::da4_7 * A[7].dy + ghrk::da4_8 * A[8].dy + ghrk::da4_9 * A[9].dy + ghrk::da4_10 * A[10].dy + ghrk::da4_11 * A[11].dy + ghrk::da4_12 * A[12].dy + ghrk::da4_13 * A[13].dy + ghrk::da4_14 * A[14].dy + ghrk::da4_15 * A[15].dy));
    const double r5 = h * (ghrk::dp5 * p0.dy + ghrk::dn5 * pn.dy + h * (ghrk::da5_1 * A[1].dy + ghrk::da5_4 * A[4].dy + ghrk::da5_5 * A[5].dy + ghrk::da5_6 * A[6].dy + ghrk::da5_7 * A[7].dy + ghrk::da5_8 * A[8].dy + ghrk::da5_9 * A[9].dy + ghrk::da5_10 * A[10].dy + ghrk::da5_11 * A[11].dy + ghrk::da5_12 * A[12].dy + ghrk::da5_13 * A[13].dy + ghrk::da5_14 * A[14].dy + ghrk::da5_15 * A[15].dy));
    const double r6 = h * (ghrk::dp6 * p0.dy + ghrk::dn6 * pn.dy + h * (ghrk::da6_1 * A[1].dy + ghrk::da6_4 * A[4].dy + ghrk::da6_5 * A[5].dy + ghrk::da6_6 * A[6].dy + ghrk::da6_7 * A[7].dy + ghrk::da6_8 * A[8].dy + ghrk::da6_9 * A[9].dy + ghrk::da6_10 * A[10].dy + ghrk::da6_11 * A[11].dy + ghrk::da6_12 * A[12].dy + ghrk::da6_13 * A[13].dy + ghrk::da6_14 * A[14].dy + ghrk::da6_15 * A[15].dy));
    const double r7 = h * (ghrk::dp7 * p0.dy + ghrk::dn7 * pn.dy + h * (ghrk::da7_1 * A[1].dy + ghrk::da7_4 * A[4].dy + ghrk::da7_5 * A[5].dy + ghrk::da7_6 * A[6].dy + ghrk::da7_7 * A[7].dy + ghrk::da7_8 * A[8].dy + ghrk::da7_9 * A[9].dy + ghrk::da7_10 * A[10].dy + ghrk::da7_11 * A[11].dy + ghrk::da7_12 * A[12].dy + ghrk::da7_13 * A[13].dy + ghrk::da7_14 * A[14].dy + ghrk::da7_15 * A[15].dy));
    yo[3] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = pn.x - p0.x;
    const double bspl = h * A[1].x - ydiff;
    const double r0 = p0.x, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * A[13].x - bspl;
    const double r4 = h * (ghdp::d41 * A[1].x + ghdp::d46 * A[6].x + ghdp::d47 * A[7].x + ghdp::d48 * A[8].x + ghdp::d49 * A[9].x + ghdp::d410 * A[10].x + ghdp::d411 * A[11].x + ghdp::d412 * A[12].x + ghdp::d413 * A[13].x + ghdp::d414 * A[14].x + ghdp::d415 * A[15].x + ghdp::d416 * A[16].x);
    const double r5 = h * (ghdp::d51 * A[1].x + ghdp::d56 * A[6].x + ghdp::d57 * A[7].x + ghdp::d58 * A[8].x + ghdp::d59 * A[9].x + ghdp::d510 * A[10].x + ghdp::d511 * A[11].x + ghdp::d512 * A[12].x + ghdp::d513 * A[13].x + ghdp::d514 * A[14].x + ghdp::d515 * A[15].x + ghdp::d516 * A[16].x);
    const double r6 = h * (ghdp::d61 * A[1].x + ghdp::d66 * A[6].x + ghdp::d67 * A[7].x + ghdp::d68 * A[8].x + ghdp::d69 * A[9].x + ghdp::d610 * A[10].x + ghdp::d611 * A[11].x + ghdp::d612 * A[12].x + ghdp::d613 * A[13].x + ghdp::d614 * A[14].x + ghdp::d615 * A[15].x + ghdp::d616 * A[16].x);
    const double r7 = h * (ghdp::d71 * A[1].x + ghdp::d76 * A[6].x + ghdp::d77 * A[7].x + ghdp::d78 * A[8].x + ghdp::d79 * A[9].x + ghdp::d710 * A[10].x + ghdp::d711 * A[11].x + ghdp::d712 * A[12].x + ghdp::d713 * A[13].x + ghdp::d714 * A[14].x + ghdp::d715 * A[15].x + ghdp::d716 * A[16].x);
    yo[4] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = pn.y - p0.y;
    const double bspl = h * A[1].y - ydiff;
    const double r0 = p0.y, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * A[13].y - bspl;
    const double r4 = h * (ghdp::d41 * A[1].y + ghdp::d46 * A[6].y + ghdp::d47 * A[7].y + ghdp::d48 * A[8].y + ghdp::d49 * A[9].y + ghdp::d410 * A[10].y + ghdp::d411 * A[11].y + ghdp::d412 * A[12].y + ghdp::d413 * A[13].y + ghdp::d414 * A[14].y + ghdp::d415 * A[15].y + ghdp::d416 * A[16].y);
    const double r5 = h * (ghdp::d51 * A[1].y + ghdp::d56 * A[6].y + ghdp::d57 * A[7].y + ghdp::d58 * A[8].y + ghdp::d59 * A[9].y + ghdp::d510 * A[10].y + ghdp::d511 * A[11].y + ghdp::d512 * A[12].y + ghdp::d513 * A[13].y + ghdp::d514 * A[14].y + ghdp::d515 * A[15].y + ghdp::d516 * A[16].y);
    const double r6 = h * (ghdp::d61 * A[1].y + ghdp::d66 * A[6].y + ghdp::d67 * A[7].y + ghdp::d68 * A[8].y + ghdp::d69 * A[9].y + ghdp::d610 * A[10].y + ghdp::d611 * A[11].y + ghdp::d612 * A[12].y + ghdp::d613 * A[13].y + ghdp::d614 * A[14].y + ghdp::d615 * A[15].y + ghdp::d616 * A[16].y);
    const double r7 = h * (ghdp::d71 * A[1].y + ghdp::d76 * A[6].y + ghdp::d77 * A[7].y + ghdp::d78 * A[8].y + ghdp::d79 * A[9].y + ghdp::d710 * A[10].y + ghdp::d711 * A[11].y + ghdp::d712 * A[12].y + ghdp::d713 * A[13].y + ghdp::d714 * A[14].y + ghdp::d715 * A[15].y + ghdp::d716 * A[16].y);
    yo[5] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = pn.dx - p0.dx;
    const double bspl = h * A[1].dx - ydiff;
    const double r0 = p0.dx, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * A[13].dx - bspl;
    const double r4 = h * (ghdp::d41 * A[1].dx + ghdp::d46 * A[6].dx + ghdp::d47 * A[7].dx + ghdp::d48 * A[8].dx + ghdp::d49 * A[9].dx + ghdp::d410 * A[10].dx + ghdp::d411 * A[11].dx + ghdp::d412 * A[12].dx + ghdp::d413 * A[13].dx + ghdp::d414 * A[14].dx + ghdp::d415 * A[15].dx + ghdp::d416 * A[16].dx);
    const double r5 = h * (ghdp::d51 * A[1].dx + ghdp::d56 * A[6].dx + ghdp::d57 * A[7].dx + ghdp::d58 * A[8].dx + ghdp::d59 * A[9].dx + ghdp::d510 * A[10].dx + ghdp::d511 * A[11].dx + ghdp::d512 * A[12].dx + ghdp::d513 * A[13].dx + ghdp::d514 * A[14].dx + ghdp::d515 * A[15].dx + ghdp::d516 * A[16].dx);
    const double r6 = h * (ghdp::d61 * A[1].dx + ghdp::d66 * A[6].dx + ghdp::d67 * A[7].dx + ghdp::d68 * A[8].dx + ghdp::d69 * A[9].dx + ghdp::d610 * A[10].dx + ghdp::d611 * A[11].dx + ghdp::d612 * A[12].dx + ghdp::d613 * A[13].dx + ghdp::d614 * A[14].dx + ghdp::d615 * A[15].dx + ghdp::d616 * A[16].dx);
    const double r7 = h * (ghdp::d71 * A[1].dx + ghdp::d76 * A[6].dx + ghdp::d77 * A[7].dx + ghdp::d78 * A[8].dx + ghdp::d79 * A[9].dx + ghdp::d710 * A[10].dx + ghdp::d711 * A[11].dx + ghdp::d712 * A[12].dx + ghdp::d713 * A[13].dx + ghdp::d714 * A[14].dx + ghdp::d715 * A[15].dx + ghdp::d716 * A[16].dx);
    yo[6] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
  {
    const double ydiff = pn.dy - p0.dy;
    const double bspl = h * A[1].dy - ydiff;
    const double r0 = p0.dy, r1 = ydiff, r2 = bspl;
    const double r3 = ydiff - h * A[13].dy - bspl;
    const double r4 = h * (ghdp::d41 * A[1].dy + ghdp::d46 * A[6].dy + ghdp::d47 * A[7].dy + ghdp::d48 * A[8].dy + ghdp::d49 * A[9].dy + ghdp::d410 * A[10].dy + ghdp::d411 * A[11].dy + ghdp::d412 * A[12].dy + ghdp::d413 * A[13].dy + ghdp::d414 * A[14].dy + ghdp::d415 * A[15].dy + ghdp::d416 * A[16].dy);
    const double r5 = h * (ghdp::d51 * A[1].dy + ghdp::d56 * A[6].dy + ghdp::d57 * A[7].dy + ghdp::d58 * A[8].dy + ghdp::d59 * A[9].dy + ghdp::d510 * A[10].dy + ghdp::d511 * A[11].dy + ghdp::d512 * A[12].dy + ghdp::d513 * A[13].dy + ghdp::d514 * A[14].dy + ghdp::d515 * A[15].dy + ghdp::d516 * A[16].dy);
    const double r6 = h * (ghdp::d61 * A[1].dy + ghdp::d66 * A[6].dy + ghdp::d67 * A[7].dy + ghdp::d68 * A[8].dy + ghdp::d69 * A[9].dy + ghdp::d610 * A[10].dy + ghdp::d611 * A[11].dy + ghdp::d612 * A[12].dy + ghdp::d613 * A[13].dy + ghdp::d614 * A[14].dy + ghdp::d615 * A[15].dy + ghdp::d616 * A[16].dy);
    const double r7 = h * (ghdp::d71 * A[1].dy + ghdp::d76 * A[6].dy + ghdp::d77 * A[7].dy + ghdp::d78 * A[8].dy + ghdp::d79 * A[9].dy + ghdp::d710 * A[10].dy + ghdp::d711 * A[11].dy + ghdp::d712 * A[12].dy + ghdp::d713 * A[13].dy + ghdp::d714 * A[14].dy + ghdp::d715 * A[15].dy + ghdp::d716 * A[16].dy);
    yo[7] = r0 + sd * (r1 + s1 * (r2 + sd * (r3 + s1 * (r4 + sd * (r5 + s1 * (r6 + sd * r7))))));
  }
}
