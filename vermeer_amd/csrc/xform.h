// 4x4 transform arithmetic shared by host and device code, written in the reference's operation
// order (every float32 rounding decides the object-space ray bit for bit).  Matrices are
// column-major, element (i, j) at m[j * 4 + i] (math/matrix4.go:50-57).
//
// A TransformDecomp key (math/animdecomp.go:13-17) is stored as kXfKeyFloats floats: T[3],
// R = (X, Y, Z, W), S[16], one pad.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

namespace vgpu {

constexpr int kXfKeyFloats = 24;

// Cofactor expansion of Matrix4Inverse / Matrix4Det (math/matrix4.go:116-180), term for term.
__host__ __device__ inline void xf_cofactors(const float* m, float* c) {
    c[0] = m[5]*m[10]*m[15] - m[5]*m[11]*m[14] - m[9]*m[6]*m[15] + m[9]*m[7]*m[14] + m[13]*m[6]*m[11] - m[13]*m[7]*m[10];
    c[4] = -m[4]*m[10]*m[15] + m[4]*m[11]*m[14] + m[8]*m[6]*m[15] - m[8]*m[7]*m[14] - m[12]*m[6]*m[11] + m[12]*m[7]*m[10];
    c[8] = m[4]*m[9]*m[15] - m[4]*m[11]*m[13] - m[8]*m[5]*m[15] + m[8]*m[7]*m[13] + m[12]*m[5]*m[11] - m[12]*m[7]*m[9];
    c[12] = -m[4]*m[9]*m[14] + m[4]*m[10]*m[13] + m[8]*m[5]*m[14] - m[8]*m[6]*m[13] - m[12]*m[5]*m[10] + m[12]*m[6]*m[9];
    c[1] = -m[1]*m[10]*m[15] + m[1]*m[11]*m[14] + m[9]*m[2]*m[15] - m[9]*m[3]*m[14] - m[13]*m[2]*m[11] + m[13]*m[3]*m[10];
    c[5] = m[0]*m[10]*m[15] - m[0]*m[11]*m[14] - m[8]*m[2]*m[15] + m[8]*m[3]*m[14] + m[12]*m[2]*m[11] - m[12]*m[3]*m[10];
    c[9] = -m[0]*m[9]*m[15] + m[0]*m[11]*m[13] + m[8]*m[1]*m[15] - m[8]*m[3]*m[13] - m[12]*m[1]*m[11] + m[12]*m[3]*m[9];
    c[13] = m[0]*m[9]*m[14] - m[0]*m[10]*m[13] - m[8]*m[1]*m[14] + m[8]*m[2]*m[13] + m[12]*m[1]*m[10] - m[12]*m[2]*m[9];
    c[2] = m[1]*m[6]*m[15] - m[1]*m[7]*m[14] - m[5]*m[2]*m[15] + m[5]*m[3]*m[14] + m[13]*m[2]*m[7] - m[13]*m[3]*m[6];
    c[6] = -m[0]*m[6]*m[15] + m[0]*m[7]*m[14] + m[4]*m[2]*m[15] - m[4]*m[3]*m[14] - m[12]*m[2]*m[7] + m[12]*m[3]*m[6];
    c[10] = m[0]*m[5]*m[15] - m[0]*m[7]*m[13] - m[4]*m[1]*m[15] + m[4]*m[3]*m[13] + m[12]*m[1]*m[7] - m[12]*m[3]*m[5];
    c[14] = -m[0]*m[5]*m[14] + m[0]*m[6]*m[13] + m[4]*m[1]*m[14] - m[4]*m[2]*m[13] - m[12]*m[1]*m[6] + m[12]*m[2]*m[5];
    c[3] = -m[1]*m[6]*m[11] + m[1]*m[7]*m[10] + m[5]*m[2]*m[11] - m[5]*m[3]*m[10] - m[9]*m[2]*m[7] + m[9]*m[3]*m[6];
    c[7] = m[0]*m[6]*m[11] - m[0]*m[7]*m[10] - m[4]*m[2]*m[11] + m[4]*m[3]*m[10] + m[8]*m[2]*m[7] - m[8]*m[3]*m[6];
    c[11] = -m[0]*m[5]*m[11] + m[0]*m[7]*m[9] + m[4]*m[1]*m[11] - m[4]*m[3]*m[9] - m[8]*m[1]*m[7] + m[8]*m[3]*m[5];
    c[15] = m[0]*m[5]*m[10] - m[0]*m[6]*m[9] - m[4]*m[1]*m[10] + m[4]*m[2]*m[9] + m[8]*m[1]*m[6] - m[8]*m[2]*m[5];
}

// Matrix4Inverse (math/matrix4.go:116-149): cofactors times 1 / det; false (out untouched) when det == 0.
__host__ __device__ inline bool xf_inverse(const float* m, float* out) {
    float c[16];
    xf_cofactors(m, c);
    float det = m[0] * c[0] + m[1] * c[4] + m[2] * c[8] + m[3] * c[12];
    if (det == 0.0f) return false;
    det = 1.0f / det;
    for (int i = 0; i < 16; i++) out[i] = c[i] * det;
    return true;
}

// Matrix4Mul (math/matrix4.go:92-102): each entry summed from +0 in k order.
__host__ __device__ inline void xf_mul(const float* a, const float* b, float* c) {
    float t[16];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            float acc = 0.0f;
            for (int k = 0; k < 4; k++) acc += a[k * 4 + i] * b[j * 4 + k];
            t[j * 4 + i] = acc;
        }
    for (int i = 0; i < 16; i++) c[i] = t[i];
}

// Matrix4MulPoint / Matrix4MulVec (math/matrix4.go:177-218)
__host__ __device__ inline void xf_point(const float* a, const float* b, float* c) {
    const float x = b[0], y = b[1], z = b[2];
    c[0] = a[0] * x + a[4] * y + a[8] * z + a[12];
    c[1] = a[1] * x + a[5] * y + a[9] * z + a[13];
    c[2] = a[2] * x + a[6] * y + a[10] * z + a[14];
}
__host__ __device__ inline void xf_vec(const float* a, const float* b, float* c) {
    const float x = b[0], y = b[1], z = b[2];
    c[0] = a[0] * x + a[4] * y + a[8] * z;
    c[1] = a[1] * x + a[5] * y + a[9] * z;
    c[2] = a[2] * x + a[6] * y + a[10] * z;
}
// Matrix4MulVec(Matrix4Transpose(a), b): the normal transform of ApplyTransform (core/shader.go:129-135)
__host__ __device__ inline void xf_vec_t(const float* a, const float* b, float* c) {
    const float x = b[0], y = b[1], z = b[2];
    c[0] = a[0] * x + a[1] * y + a[2] * z;
    c[1] = a[4] * x + a[5] * y + a[6] * z;
    c[2] = a[8] * x + a[9] * y + a[10] * z;
}

// TransformDecompLerp then TransformDecompToMatrix4 (math/animdecomp.go:66-80) of key records a and b
// at t: T lerped, R by QuatSlerp (math/quat.go:30-67, its trigonometry in float64 as Go's math
// package computes it), S lerped entry by entry, M = Translate(T) (QuatToMatrix4(R) S)
// (math/quat.go:80-104).  With a == b and t == 0 this is TransformDecompToMatrix4 of the key alone.
__host__ __device__ inline void xf_decomp_matrix(const float* a, const float* b, float t, float* M) {
    float T[3], q[4];
    for (int i = 0; i < 3; i++) T[i] = (1.0f - t) * a[i] + t * b[i];
    const float* qa = a + 3;
    const float* qb = b + 3;
    const float c = qa[3] * qb[3] + qa[0] * qb[0] + qa[1] * qb[1] + qa[2] * qb[2];
    if (fabsf(c) >= 1.0f) {
        for (int i = 0; i < 4; i++) q[i] = qa[i];
    } else {
        const float half = (float)acos((double)c);
        const float sh = sqrtf(1.0f - c * c);
        if (fabsf(sh) < 0.001f) {
            for (int i = 0; i < 4; i++) q[i] = qa[i] * 0.5f + qb[i] * 0.5f;
        } else {
            const float ra = (float)sin((double)((1 - t) * half)) / sh;
            const float rb = (float)sin((double)(t * half)) / sh;
            for (int i = 0; i < 4; i++) q[i] = qa[i] * ra + qb[i] * rb;
        }
    }
    const float n = 1.0f / sqrtf(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    const float X = q[0] * n, Y = q[1] * n, Z = q[2] * n, W = q[3] * n;
    float R[16] = {0};
    R[0] = 1 - 2 * Y * Y - 2 * Z * Z; R[4] = 2 * X * Y - 2 * W * Z; R[8] = 2 * X * Z + 2 * W * Y;
    R[1] = 2 * X * Y + 2 * W * Z; R[5] = 1 - 2 * X * X - 2 * Z * Z; R[9] = 2 * Y * Z - 2 * W * X;
    R[2] = 2 * X * Z - 2 * W * Y; R[6] = 2 * Y * Z + 2 * W * X; R[10] = 1 - 2 * X * X - 2 * Y * Y;
    R[15] = 1.0f;
    float S[16];
    for (int i = 0; i < 16; i++) S[i] = (1.0f - t) * a[7 + i] + t * b[7 + i];
    float RS[16];
    xf_mul(R, S, RS);
    float Tm[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, T[0], T[1], T[2], 1};
    xf_mul(Tm, RS, M);
}

}  // namespace vgpu
