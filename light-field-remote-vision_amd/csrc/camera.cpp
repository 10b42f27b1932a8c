// Pinhole camera decomposition (host, FP64).
#include <cmath>
#include <cstring>

namespace lfrv {

static void mul3(const double *A, const double *B, double *C)
{
    double T[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += A[3 * i + k] * B[3 * k + j];
            T[3 * i + j] = s;
        }
    std::memcpy(C, T, sizeof T);
}

// P = K [R | t] from a row-major 3x4 P: RQ decomposition by three Givens
// rotations, positive diagonal, det R = +1, K(2,2) = 1 (the algorithm of
// sources/pinholeCamera.cpp:17-94, in double precision).  C = -R^T t.
void decompose_P(const double *P, double *K, double *R, double *t, double *C)
{
    double k[9], Q[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) k[3 * i + j] = P[4 * i + j];
    if (k[7] != 0) {  // zero K(2,1)
        double c = -k[8], s = k[7];
        const double l = std::sqrt(c * c + s * s);
        c /= l; s /= l;
        const double G[9] = {1, 0, 0, 0, c, -s, 0, s, c}, GT[9] = {1, 0, 0, 0, c, s, 0, -s, c};
        mul3(k, G, k);
        mul3(GT, Q, Q);
    }
    if (k[6] != 0) {  // zero K(2,0)
        double c = k[8], s = k[6];
        const double l = std::sqrt(c * c + s * s);
        c /= l; s /= l;
        const double G[9] = {c, 0, s, 0, 1, 0, -s, 0, c}, GT[9] = {c, 0, -s, 0, 1, 0, s, 0, c};
        mul3(k, G, k);
        mul3(GT, Q, Q);
    }
    if (k[3] != 0) {  // zero K(1,0)
        double c = -k[4], s = k[3];
        const double l = std::sqrt(c * c + s * s);
        c /= l; s /= l;
        const double G[9] = {c, -s, 0, s, c, 0, 0, 0, 1}, GT[9] = {c, s, 0, -s, c, 0, 0, 0, 1};
        mul3(k, G, k);
        mul3(GT, Q, Q);
    }
    double r[9];
    std::memcpy(r, Q, sizeof r);
    if (k[8] < 0)
        for (int i = 0; i < 9; i++) { k[i] = -k[i]; r[i] = -r[i]; }
    if (k[4] < 0) {
        for (int i = 0; i < 3; i++) k[3 * i + 1] = -k[3 * i + 1];
        for (int j = 0; j < 3; j++) r[3 + j] = -r[3 + j];
    }
    if (k[0] < 0) {
        for (int i = 0; i < 3; i++) k[3 * i] = -k[3 * i];
        for (int j = 0; j < 3; j++) r[j] = -r[j];
    }
    {   // t = K^-1 P[:,3], K upper triangular
        const double z = P[11] / k[8], y = (P[7] - k[5] * z) / k[4], x = (P[3] - k[1] * y - k[2] * z) / k[0];
        t[0] = x; t[1] = y; t[2] = z;
    }
    const double det = r[0] * (r[4] * r[8] - r[5] * r[7]) - r[1] * (r[3] * r[8] - r[5] * r[6]) +
                       r[2] * (r[3] * r[7] - r[4] * r[6]);
    if (det < 0) {
        for (int i = 0; i < 9; i++) r[i] = -r[i];
        for (int i = 0; i < 3; i++) t[i] = -t[i];
    }
    const double s22 = k[8];
    for (int i = 0; i < 9; i++) k[i] /= s22;
    std::memcpy(K, k, sizeof k);
    std::memcpy(R, r, sizeof r);
    for (int i = 0; i < 3; i++) C[i] = -(r[i] * t[0] + r[3 + i] * t[1] + r[6 + i] * t[2]);
}

}  // namespace lfrv
