// Host-side mesh index construction and geometry (FP64).
//
// Replaces, in one call, what adFVM assembles across Python and C++:
//   Mesh.populateSizes            adFVM/mesh.py:517-523
//   cmesh.buildBeforeWrite        adFVM/cpp/cmesh.cpp:3-63
//   Mesh.splitPatches             adFVM/mesh.py:238-249 (cellStartFace)
//   cmesh.build geometry          adFVM/cpp/cmesh.cpp:65-165
//   Mesh.createGhostCells         adFVM/mesh.py:758-819 (single rank)
//   cmesh.build face weights      adFVM/cpp/cmesh.cpp:173-238
//   cellOwner / cellNeighbours    adFVM/cpp/cmesh.cpp:240-269
//   volumesL / volumesR           adFVM/mesh.py:108-109
// Compiled with -ffp-contract=off so every floating-point expression is
// evaluated in the written order, as numpy evaluates the oracle's.
#include <cmath>
#include <cstring>
#include <vector>
#include <string>

#include "adfvm.h"

namespace {
thread_local std::string g_mesh_error;

int fail(const std::string& msg) {
    g_mesh_error = msg;
    return ADFVM_ERR_INVALID;
}

inline void cross(const double* a, const double* b, double* out) {
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}
}  // namespace

extern "C" const char* adfvm_mesh_last_error(void) { return g_mesh_error.c_str(); }

extern "C" int adfvm_mesh_sizes(const adfvm_mesh_input* in, adfvm_mesh_sizes_t* out) {
    if (!in || !out) return fail("null argument");
    if (in->nFaces <= 0 || in->nInternalFaces < 0 || in->nInternalFaces > in->nFaces)
        return fail("bad face counts");
    int32_t maxOwner = -1;
    for (int32_t f = 0; f < in->nFaces; f++) {
        int32_t o = in->owner[f];
        if (o < 0) return fail("negative owner");
        if (o > maxOwner) maxOwner = o;
    }
    out->nInternalFaces = in->nInternalFaces;
    out->nFaces = in->nFaces;
    out->nBoundaryFaces = in->nFaces - in->nInternalFaces;
    out->nInternalCells = maxOwner + 1;
    out->nGhostCells = out->nBoundaryFaces;
    out->nCells = out->nInternalCells + out->nGhostCells;
    return ADFVM_OK;
}

extern "C" int adfvm_mesh_build(const adfvm_mesh_input* in, adfvm_mesh_output* out) {
    adfvm_mesh_sizes_t sz;
    int rc = adfvm_mesh_sizes(in, &sz);
    if (rc != ADFVM_OK) return rc;
    const int32_t nF = sz.nFaces, nIF = sz.nInternalFaces, nIC = sz.nInternalCells;
    const int32_t nBF = sz.nBoundaryFaces;
    const double* points = in->points;
    const int32_t* faces = in->faces;
    const int32_t* owner = in->owner;
    for (int32_t f = 0; f < nIF; f++) {
        int32_t n = in->neighbour[f];
        if (n < 0 || n >= nIC) return fail("internal neighbour out of range");
    }
    for (int32_t f = 0; f < nF * 4; f++)
        if (faces[f] < 0 || faces[f] >= in->nPoints) return fail("face point index out of range");
    // patches must tile [nIF, nF)
    {
        std::vector<int> covered(nBF, 0);
        for (int32_t p = 0; p < in->nPatches; p++) {
            int32_t s = in->patchStartFace[p], n = in->patchNFaces[p];
            if (s < nIF || n < 0 || s + n > nF) return fail("patch range outside boundary faces");
            for (int32_t f = s; f < s + n; f++) covered[f - nIF]++;
            if (in->patchType[p] == ADFVM_PATCH_CYCLIC) {
                int32_t q = in->patchNeighbour[p];
                if (q < 0 || q >= in->nPatches || in->patchNFaces[q] != n)
                    return fail("cyclic patch without a matching neighbourPatch");
            }
        }
        for (int32_t i = 0; i < nBF; i++)
            if (covered[i] != 1) return fail("patches do not tile the boundary faces");
    }

    // ---- cellFaces: owner faces in face order, then neighbour faces (cmesh.cpp:9-19)
    int32_t* cellFaces = out->cellFaces;
    {
        std::vector<int32_t> fill(nIC, 0);
        for (int32_t f = 0; f < nF; f++) {
            int32_t p = owner[f];
            if (fill[p] >= 6) return fail("cell with more than 6 faces (hex meshes only)");
            cellFaces[p * 6 + fill[p]++] = f;
        }
        for (int32_t f = 0; f < nIF; f++) {
            int32_t p = in->neighbour[f];
            if (fill[p] >= 6) return fail("cell with more than 6 faces (hex meshes only)");
            cellFaces[p * 6 + fill[p]++] = f;
        }
        for (int32_t c = 0; c < nIC; c++)
            if (fill[c] != 6) return fail("cell with fewer than 6 faces (hex meshes only)");
    }

    // ---- cells: hex point lists (cmesh.cpp:20-60; cyclic predecessor (k+3)%4)
    if (out->cells) {
        for (int32_t i = 0; i < nIC; i++) {
            const int32_t* first = faces + 4 * cellFaces[i * 6];
            int32_t* cp = out->cells + 8 * i;
            for (int j = 0; j < 4; j++) {
                int32_t point = first[j];
                bool found = false;
                for (int n = 1; n < 6 && !found; n++) {
                    const int32_t* nxt = faces + 4 * cellFaces[i * 6 + n];
                    for (int k = 0; k < 4; k++) {
                        if (nxt[k] == point) {
                            int l = (k + 1) % 4;
                            for (int m = 0; m < 4; m++)
                                if (first[m] == nxt[l]) { l = (k + 3) % 4; break; }
                            cp[4 + j] = nxt[l];
                            found = true;
                            break;
                        }
                    }
                }
            }
            for (int j = 0; j < 4; j++) cp[j] = first[j];
        }
    }

    // ---- face normals, centres, areas (cmesh.cpp:70-121)
    double* normals = out->normals;
    double* faceCentres = out->faceCentres;
    double* areas = out->areas;
    for (int32_t f = 0; f < nF; f++) {
        const double* a = points + 3 * faces[4 * f + 0];
        const double* b = points + 3 * faces[4 * f + 1];
        const double* c = points + 3 * faces[4 * f + 2];
        double v1[3], v2[3], nrm[3];
        for (int i = 0; i < 3; i++) { v1[i] = a[i] - b[i]; v2[i] = b[i] - c[i]; }
        cross(v1, v2, nrm);
        double Ns = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
        for (int i = 0; i < 3; i++) normals[3 * f + i] = nrm[i] / Ns;
        double fc[3] = {0, 0, 0};
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 4; j++) fc[i] += points[3 * faces[4 * f + j] + i];
            fc[i] /= 4;
        }
        double area = 0, sc[3] = {0, 0, 0};
        for (int j = 0; j < 4; j++) {
            const double* p = points + 3 * faces[4 * f + j];
            const double* q = points + 3 * faces[4 * f + (j + 1) % 4];
            double avg[3], w1[3], w2[3], N[3];
            for (int i = 0; i < 3; i++) {
                avg[i] = (p[i] + q[i] + fc[i]) / 3;
                w1[i] = q[i] - p[i];
                w2[i] = fc[i] - p[i];
            }
            cross(w1, w2, N);
            double ns = std::sqrt(N[0] * N[0] + N[1] * N[1] + N[2] * N[2]);
            area += ns / 2;
            for (int i = 0; i < 3; i++) sc[i] += ns * avg[i] / 2;
        }
        areas[f] = area;
        for (int i = 0; i < 3; i++) faceCentres[3 * f + i] = sc[i] / area;
    }

    // ---- cell centres and volumes (cmesh.cpp:126-163)
    double* cellCentres = out->cellCentres;   // nCells x 3
    double* volumes = out->volumes;
    for (int32_t c = 0; c < nIC; c++) {
        double cc[3] = {0, 0, 0};
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 6; j++) cc[i] += faceCentres[3 * cellFaces[6 * c + j] + i];
            cc[i] /= 6;
        }
        double vol = 0, sc[3] = {0, 0, 0};
        for (int j = 0; j < 6; j++) {
            int32_t f = cellFaces[6 * c + j];
            const double* fcj = faceCentres + 3 * f;
            double area = areas[f];
            const double* N = normals + 3 * f;
            double v = 0, avgc[3];
            for (int i = 0; i < 3; i++) {
                double height = cc[i] - fcj[i];
                double areaN = area * N[i];
                v += areaN * height;
                avgc[i] = 3. / 4 * fcj[i] + 1. / 4 * cc[i];
            }
            v = std::fabs(v / 3);
            vol += v;
            for (int i = 0; i < 3; i++) sc[i] += v * avgc[i];
        }
        volumes[c] = vol;
        for (int i = 0; i < 3; i++) cellCentres[3 * c + i] = sc[i] / vol;
    }

    // ---- ghost cells (mesh.py:758-819)
    int32_t* neighbour = out->neighbour;   // nFaces, ghost-extended
    for (int32_t f = 0; f < nIF; f++) neighbour[f] = in->neighbour[f];
    int32_t nLocal = nIC;
    for (int32_t p = 0; p < in->nPatches; p++) {
        int32_t start = in->patchStartFace[p], n = in->patchNFaces[p];
        int32_t cstart = start - nIF + nIC;
        if (n == 0) continue;
        if (in->patchType[p] != ADFVM_PATCH_PROCESSOR) nLocal += n;
        for (int32_t k = 0; k < n; k++) neighbour[start + k] = cstart + k;
        if (in->patchType[p] == ADFVM_PATCH_CYCLIC) {
            int32_t nstart = in->patchStartFace[in->patchNeighbour[p]];
            double tr[3];
            for (int i = 0; i < 3; i++) tr[i] = faceCentres[3 * start + i] - faceCentres[3 * nstart + i];
            for (int32_t k = 0; k < n; k++) {
                int32_t idx = owner[nstart + k];
                for (int i = 0; i < 3; i++) cellCentres[3 * (cstart + k) + i] = tr[i] + cellCentres[3 * idx + i];
            }
        } else {
            for (int32_t k = 0; k < n; k++)
                for (int i = 0; i < 3; i++) cellCentres[3 * (cstart + k) + i] = faceCentres[3 * (start + k) + i];
        }
    }
    if (out->nLocalCells) *out->nLocalCells = nLocal;

    // ---- deltas, weights, linear/quadratic weights (cmesh.cpp:173-238)
    for (int32_t f = 0; f < nF; f++) {
        const double* P = cellCentres + 3 * owner[f];
        const double* N = cellCentres + 3 * neighbour[f];
        double delta[3];
        for (int i = 0; i < 3; i++) delta[i] = P[i] - N[i];
        double d = std::sqrt(delta[0] * delta[0] + delta[1] * delta[1] + delta[2] * delta[2]);
        out->deltas[f] = d;
        for (int i = 0; i < 3; i++) out->deltasUnit[3 * f + i] = -delta[i] / d;
        const double* F = faceCentres + 3 * f;
        const double* nrm = normals + 3 * f;
        double nD = 0, pD = 0, nFv[3], pFv[3];
        for (int i = 0; i < 3; i++) {
            nFv[i] = F[i] - N[i];
            pFv[i] = F[i] - P[i];
            nD += nFv[i] * nrm[i];
            pD += pFv[i] * nrm[i];
        }
        nD = std::fabs(nD);
        pD = std::fabs(pD);
        out->weights[f] = nD / (nD + pD);
        double w1 = 0, w2 = 0;
        d = 0;
        for (int i = 0; i < 3; i++) {
            w1 += -delta[i] * pFv[i];
            w2 += delta[i] * nFv[i];
            d += delta[i] * delta[i];
        }
        w1 /= d;
        w2 /= d;
        out->linearWeights[2 * f + 0] = w1 / 3;
        out->linearWeights[2 * f + 1] = w2 / 3;
        for (int i = 0; i < 3; i++) {
            out->quadraticWeights[6 * f + i] = 2. / 3 * pFv[i] + 1. / 3 * (pFv[i] + w1 * delta[i]);
            out->quadraticWeights[6 * f + 3 + i] = 2. / 3 * nFv[i] + 1. / 3 * (nFv[i] - w2 * delta[i]);
        }
    }

    // ---- cellOwner / cellNeighbours (cmesh.cpp:240-269)
    for (int32_t c = 0; c < nIC; c++) {
        for (int j = 0; j < 6; j++) {
            int32_t f = cellFaces[6 * c + j];
            int32_t p = owner[f], n = neighbour[f];
            if (p != c) {
                out->cellOwner[6 * c + j] = 0;
                out->cellNeighbours[6 * c + j] = p;
                if (out->cellNeighboursMatOp) out->cellNeighboursMatOp[6 * c + j] = p;
            } else {
                out->cellOwner[6 * c + j] = 1;
                out->cellNeighbours[6 * c + j] = n;
                if (out->cellNeighboursMatOp) out->cellNeighboursMatOp[6 * c + j] = n < nIC ? n : -1;
            }
        }
    }
    for (int32_t f = 0; f < nF; f++) out->volumesL[f] = volumes[owner[f]];
    for (int32_t f = 0; f < nIF; f++) out->volumesR[f] = volumes[neighbour[f]];
    return ADFVM_OK;
}
