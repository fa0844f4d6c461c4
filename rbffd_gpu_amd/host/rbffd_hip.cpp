#include "rbffd_hip.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>

namespace rbfx {

namespace {
[[noreturn]] void die(rbfx_ctx* ctx, const char* what, int rc) {
    std::cout << "[RBFFD_HIP] ERROR in " << what << " (" << rc << "): " << rbfx_last_error(ctx) << std::endl;
    std::exit(EXIT_FAILURE);
}
#define RBFX_CALL(ctx, call)                 \
    do {                                     \
        int rc_ = (call);                    \
        if (rc_ != RBFX_OK) die(ctx, #call, rc_); \
    } while (0)

const uint32_t kSolvable = RBFFD::X | RBFFD::Y | RBFFD::Z | RBFFD::LAPL | RBFFD::R | RBFFD::HV | RBFFD::LAMBDA |
                           RBFFD::THETA | RBFFD::LSFC | RBFFD::XSFC | RBFFD::YSFC | RBFFD::ZSFC | RBFFD::INTERP;
const uint32_t kAlt = RBFFD::ALT_XSFC | RBFFD::ALT_YSFC | RBFFD::ALT_ZSFC;
}  // namespace

// ---- RBFFD ------------------------------------------------------------------------------------

RBFFD::RBFFD(DerTypes typesToCompute, Grid* grid, int dim_num_, int rank_)
    : computedTypes(typesToCompute), grid_ref(*grid), dim_num(dim_num_), rank(rank_) {
    unsigned int nb = grid_ref.getStencilsSize();
    for (int i = 0; i < NUM_DERIVATIVE_TYPES; i++) weights[i].assign(nb, nullptr);
    var_epsilon.assign(nb, 0.0);
}

RBFFD::~RBFFD() {
    for (int j = 0; j < NUM_DERIVATIVE_TYPES; j++)
        for (double* p : weights[j]) delete[] p;
}

void RBFFD::setUseHyperviscosity(int tf) {
    useHyperviscosity = tf;
    if (tf) appendDerTypes(HV);
    else removeDerType(HV);
}

double RBFFD::getHVScalar() const { return -hv_gamma * std::pow((double)grid_ref.getGlobalNodeListSize(), -hv_k); }

void RBFFD::setWeightType(WeightType t) {
    if (t != Direct) {
        std::cout << "[RBFFD] ERROR: only the Direct weight method is provided" << std::endl;
        std::exit(EXIT_FAILURE);
    }
}

void RBFFD::setEpsilonByParameters(double c1, double c2) {
    setEpsilon(c1 * std::sqrt((double)grid_ref.getGlobalNodeListSize()) - c2);
}

void RBFFD::setEpsilonByStencilSize() {
    // rbffd.h:300-349 (values tuned for the cosine bell)
    double c1, c2;
    switch (grid_ref.getMaxStencilSize()) {
        case 17: c1 = 0.026; c2 = 0.08; hv_k = 2; hv_gamma = 8e-4; break;
        case 31: c1 = 0.035; c2 = 0.1; hv_k = 4; hv_gamma = 5e-2; break;
        case 50: c1 = 0.044; c2 = 0.14; hv_k = 6; hv_gamma = 5e-1; break;
        case 101: c1 = 0.058; c2 = 0.16; hv_k = 8; hv_gamma = 5e-2; break;
        default:
            std::cout << "[RBFFD] Error: setEpsilonByStencilSize does not support stencil size "
                      << grid_ref.getMaxStencilSize() << " at this time. Try using 17, 31, 50, and 101\n";
            std::exit(EXIT_FAILURE);
    }
    setEpsilonByParameters(c1, c2);
}

void RBFFD::setEpsilon(double eps) {
    for (double& e : var_epsilon) e = eps;
}

void RBFFD::setEpsilon(std::vector<double>& lst) {
    if (lst.size() != var_epsilon.size()) {
        std::cout << "ERROR! length of support params list does not match" << std::endl;
        std::exit(EXIT_FAILURE);
    }
    var_epsilon = lst;
}

void RBFFD::storeHostWeights(const std::vector<double>& w, const std::vector<int32_t>& off, DerTypes types) {
    size_t nnz = (size_t)off.back();
    int t = 0;
    for (int b = 0; b < 16; b++) {
        if (!(types & (1u << b))) continue;
        for (size_t s = 0; s + 1 < off.size(); s++) {
            int n = off[s + 1] - off[s];
            delete[] weights[b][s];
            weights[b][s] = new double[n];
            std::memcpy(weights[b][s], &w[(size_t)t * nnz + off[s]], sizeof(double) * n);
        }
        t++;
    }
}

void RBFFD::writeToFile(DerType which, const std::string& filename) {
    std::vector<StencilType>& st = grid_ref.getStencils();
    std::vector<double*>& w = weights[getDerTypeIndx(which)];
    unsigned long nz = 0;
    for (auto& s : st) nz += s.size();
    FILE* f = std::fopen(filename.c_str(), "w");
    if (!f) {
        std::cout << "[RBFFD] cannot open " << filename << std::endl;
        std::exit(EXIT_FAILURE);
    }
    std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
    std::fprintf(f, "%u %u %lu\n", grid_ref.getStencilsSize(), grid_ref.getNodeListSize(), nz);
    for (size_t i = 0; i < st.size(); i++)
        for (size_t j = 0; j < st[i].size(); j++) {
            int ii = st[i][0] + 1, jj = st[i][j] + 1;
            double kk = w[i][j];
            if (asciiWeights) {
                std::fprintf(f, "%d %d %24.16le\n", ii, jj, kk);
            } else {
                std::fwrite(&ii, sizeof(int), 1, f);
                std::fwrite(&jj, sizeof(int), 1, f);
                std::fwrite(&kk, sizeof(double), 1, f);
            }
        }
    std::fclose(f);
}

int RBFFD::loadFromFile(DerType which, const std::string& filename) {
    // RBFFD::loadFromAsciiFile / loadFromBinaryFile (rbffd.cpp:981-1160): the banner, the
    // size line "M N nz", then nz (row, col, weight) triples, 1-based, as text or binary.
    // Return codes as there: 1 no file, 2 bad banner, 4 bad size line, 5 M or nz differ from
    // the grid's stencils; triples fill weights[row][j] in file order (j restarts at each new
    // row), the column index is not used.
    FILE* f = std::fopen(filename.c_str(), "r");
    if (!f) {
        std::cout << "File not found: " << filename << std::endl;
        return 1;
    }
    char line[1024];
    if (!std::fgets(line, sizeof(line), f) || std::strncmp(line, "%%MatrixMarket", 14) != 0) {
        std::cout << "Could not process MatrixMarket Banner in " << filename << std::endl;
        std::fclose(f);
        return 2;
    }
    long M = 0, N = 0, nz = 0;
    do {
        if (!std::fgets(line, sizeof(line), f)) {
            std::fclose(f);
            return 4;
        }
    } while (line[0] == '%');
    if (std::sscanf(line, "%ld %ld %ld", &M, &N, &nz) != 3) {
        std::cout << "Error! failed to parse file contents" << std::endl;
        std::fclose(f);
        return 4;
    }
    std::vector<StencilType>& st = grid_ref.getStencils();
    std::vector<double*>& w = weights[getDerTypeIndx(which)];
    long expected_nz = 0;
    for (auto& s : st) expected_nz += (long)s.size();
    if (M != (long)w.size()) {
        std::cout << "Error! not enough stencils in the file" << std::endl;
        std::fclose(f);
        return 5;
    }
    if (nz != expected_nz) {
        std::cout << "Error! not enough weights in the file" << std::endl;
        std::fclose(f);
        return 5;
    }
    std::vector<int> I(nz);
    std::vector<double> val(nz);
    for (long q = 0; q < nz; q++) {
        int ii = 0, jj = 0;
        double kk = 0.0;
        bool ok = asciiWeights ? std::fscanf(f, "%d %d %le", &ii, &jj, &kk) == 3
                               : (std::fread(&ii, sizeof(int), 1, f) == 1 && std::fread(&jj, sizeof(int), 1, f) == 1 &&
                                  std::fread(&kk, sizeof(double), 1, f) == 1);
        if (!ok || ii < 1 || ii > M) {
            std::cout << "Error! failed to parse file contents" << std::endl;
            std::fclose(f);
            return 4;
        }
        I[q] = ii - 1;
        val[q] = kk;
    }
    std::fclose(f);
    for (size_t i = 0; i < st.size(); i++) {
        delete[] w[i];
        w[i] = new double[st[i].size()]();
    }
    long j = 0;
    int row = 0;
    for (long q = 0; q < nz; q++) {
        if (I[q] != row) {
            j = 0;
            row = I[q];
        }
        if (j >= (long)st[row].size()) {
            std::printf("[RBFFD] ERROR! File contains too many nonzeros. I have not allocated enough memory!\n");
            std::exit(EXIT_FAILURE);
        }
        w[row][j++] = val[q];
    }
    computedTypes |= which;
    weightsModified = true;
    return 0;
}

// ---- RBFFD_HIP -------------------------------------------------------------------------------

RBFFD_HIP::RBFFD_HIP(DerTypes typesToCompute, Grid* grid, int dim_num_, int rank_)
    : RBFFD(typesToCompute, grid, dim_num_, rank_) {
    int rc = rbfx_create(0, &ctx);
    if (rc != RBFX_OK) {
        std::cout << "[RBFFD_HIP] ERROR: " << rbfx_last_error(nullptr) << std::endl;
        std::exit(EXIT_FAILURE);
    }
}

RBFFD_HIP::~RBFFD_HIP() {
    for (rbfx_ell* e : ell_by_type) rbfx_ell_destroy(e);
    for (auto& p : fused) rbfx_ell_destroy(p.second);
    for (void* p : {(void*)d_xyz, (void*)d_idx, (void*)d_off, (void*)d_csr, (void*)d_w, (void*)d_wpad, d_u, d_out})
        rbfx_dev_free(ctx, p);
    rbfx_destroy(ctx);
}

void RBFFD_HIP::uploadGeometry() {
    std::vector<NodeType>& nodes = grid_ref.getNodeList();
    std::vector<StencilType>& st = grid_ref.getStencils();
    size_t N = nodes.size(), S = st.size();
    std::vector<double> xyz(3 * N);
    for (size_t i = 0; i < N; i++)
        for (int d = 0; d < 3; d++) xyz[3 * i + d] = nodes[i][d];
    st_off.assign(S + 1, 0);
    width = 0;
    for (size_t s = 0; s < S; s++) {
        st_off[s + 1] = st_off[s] + (int32_t)st[s].size();
        width = st[s].size() > width ? (unsigned)st[s].size() : width;
    }
    std::vector<int32_t> csr(st_off[S]), pad(S * (size_t)width);
    for (size_t s = 0; s < S; s++)
        for (size_t j = 0; j < width; j++) {
            int v = j < st[s].size() ? st[s][j] : st[s][0];  // pad with the centre (rbffd_cl.cpp:238-242)
            if (j < st[s].size()) csr[st_off[s] + j] = v;
            pad[s * width + j] = v;
        }
    for (void* p : {(void*)d_xyz, (void*)d_idx, (void*)d_off, (void*)d_csr}) rbfx_dev_free(ctx, p);
    RBFX_CALL(ctx, rbfx_dev_alloc(ctx, xyz.size() * 8, (void**)&d_xyz));
    RBFX_CALL(ctx, rbfx_dev_alloc(ctx, pad.size() * 4, (void**)&d_idx));
    RBFX_CALL(ctx, rbfx_dev_alloc(ctx, st_off.size() * 4, (void**)&d_off));
    RBFX_CALL(ctx, rbfx_dev_alloc(ctx, csr.size() * 4, (void**)&d_csr));
    RBFX_CALL(ctx, rbfx_h2d(ctx, d_xyz, xyz.data(), xyz.size() * 8));
    RBFX_CALL(ctx, rbfx_h2d(ctx, d_idx, pad.data(), pad.size() * 4));
    RBFX_CALL(ctx, rbfx_h2d(ctx, d_off, st_off.data(), st_off.size() * 4));
    RBFX_CALL(ctx, rbfx_h2d(ctx, d_csr, csr.data(), csr.size() * 4));
}

void RBFFD_HIP::computeAllWeightsForAllStencils() {
    uploadGeometry();
    size_t S = grid_ref.getStencilsSize();
    uint32_t solve_types = computedTypes & kSolvable;
    if (computedTypes & kAlt) solve_types |= X | Y | Z;
    int T = __builtin_popcount(solve_types);
    size_t nnz = (size_t)st_off[S];
    double* d_eps = nullptr;
    double* d_diag = nullptr;
    rbfx_dev_free(ctx, d_w);
    RBFX_CALL(ctx, rbfx_dev_alloc(ctx, (size_t)T * nnz * 8, (void**)&d_w));
    RBFX_CALL(ctx, rbfx_dev_alloc(ctx, S * 8, (void**)&d_eps));
    RBFX_CALL(ctx, rbfx_dev_alloc(ctx, S * 4 * 8, (void**)&d_diag));
    RBFX_CALL(ctx, rbfx_h2d(ctx, d_eps, var_epsilon.data(), S * 8));
    int32_t nsing = 0;
    int rc = rbfx_weights(ctx, (int32_t)S, d_off, d_csr, (int32_t)grid_ref.getNodeListSize(), d_xyz, d_eps, solve_types,
                          dim_num, poly_deg, (int32_t)hv_k, getHVScalar(), solve_mode, d_w, d_diag, &nsing);
    if (rc != RBFX_OK) die(ctx, "rbfx_weights", rc);
    std::vector<double> w((size_t)T * nnz);
    diag.resize(S * 4);
    RBFX_CALL(ctx, rbfx_d2h(ctx, w.data(), d_w, w.size() * 8));
    RBFX_CALL(ctx, rbfx_d2h(ctx, diag.data(), d_diag, diag.size() * 8));
    rbfx_dev_free(ctx, d_eps);
    rbfx_dev_free(ctx, d_diag);
    storeHostWeights(w, st_off, solve_types);
    if (computedTypes & kAlt) {
        // ALT_*SFC = (I - x x^T) [X, Y, Z] (rbffd.cpp:337-347)
        std::vector<StencilType>& st = grid_ref.getStencils();
        for (size_t i = 0; i < S; i++) {
            NodeType c = grid_ref.getNode(i);
            double xx = c.x(), yy = c.y(), zz = c.z();
            size_t n = st[i].size();
            for (int a = 0; a < 3; a++) {
                int idx = ALT_XSFC_i + a;
                if (!(computedTypes & (ALT_XSFC << a))) continue;
                delete[] weights[idx][i];
                weights[idx][i] = new double[n];
                double r[3][3] = {{1 - xx * xx, -xx * yy, -xx * zz}, {-xx * yy, 1 - yy * yy, -yy * zz}, {-xx * zz, -yy * zz, 1 - zz * zz}};
                for (size_t j = 0; j < n; j++)
                    weights[idx][i][j] = r[a][0] * weights[X_i][i][j] + r[a][1] * weights[Y_i][i][j] + r[a][2] * weights[Z_i][i][j];
            }
        }
    }
    weightsModified = true;
}

void RBFFD_HIP::setGPUPrecision(int storage_, int accum_) {
    storage = storage_;
    accum = accum_;
    weightsModified = true;
}

void RBFFD_HIP::updateWeightsOnGPU(bool) {
    if (!weightsModified) return;
    for (rbfx_ell*& e : ell_by_type) {
        rbfx_ell_destroy(e);
        e = nullptr;
    }
    for (auto& p : fused) rbfx_ell_destroy(p.second);
    fused.clear();
    if (!d_idx) uploadGeometry();
    weightsModified = false;
}

rbfx_ell* RBFFD_HIP::getGPUOperatorSet(DerTypes types) { return getGPUOperatorSet(types, storage); }

rbfx_ell* RBFFD_HIP::getGPUOperatorSet(DerTypes types, int storage_) {
    updateWeightsOnGPU(true);
    const uint64_t key = (uint64_t)types | ((uint64_t)(uint32_t)storage_ << 32);
    for (auto& p : fused)
        if (p.first == key) return p.second;
    rbfx_ell* e = buildEll(types, storage_);
    fused.push_back({key, e});
    return e;
}

rbfx_ell* RBFFD_HIP::buildEll(DerTypes types, int storage_) {
    size_t S = grid_ref.getStencilsSize();
    std::vector<StencilType>& st = grid_ref.getStencils();
    std::vector<const double*> ptrs;
    std::vector<uint32_t> bits;
    std::vector<double*> owned;
    for (int b = 0; b < 16; b++) {
        if (!(types & (1u << b))) continue;
        if (!weights[b].size() || !weights[b][0]) {
            std::cout << "[RBFFD_HIP] ERROR: weights of type " << (1u << b) << " were not computed" << std::endl;
            std::exit(EXIT_FAILURE);
        }
        std::vector<double> pad(S * (size_t)width, 0.0);
        for (size_t s = 0; s < S; s++) std::memcpy(&pad[s * width], weights[b][s], sizeof(double) * st[s].size());
        double* d = nullptr;
        RBFX_CALL(ctx, rbfx_dev_alloc(ctx, pad.size() * 8, (void**)&d));
        RBFX_CALL(ctx, rbfx_h2d(ctx, d, pad.data(), pad.size() * 8));
        ptrs.push_back(d);
        owned.push_back(d);
        bits.push_back(1u << b);
    }
    rbfx_ell* e = nullptr;
    int rc = rbfx_ell_create(ctx, (int32_t)S, (int32_t)grid_ref.getNodeListSize(), (int32_t)width, d_idx, (int32_t)ptrs.size(),
                             bits.data(), ptrs.data(), storage_, &e);
    for (double* d : owned) rbfx_dev_free(ctx, d);
    if (rc != RBFX_OK) die(ctx, "rbfx_ell_create", rc);
    // rows (stencil centres) processed along a Morton curve of their nodes: neighbouring
    // stencils run side by side and share their gathers in cache; results are unchanged
    std::vector<double> cxyz(3 * S);
    for (size_t s = 0; s < S; s++) {
        const NodeType& p = grid_ref.getNode((int)s);
        cxyz[3 * s] = p.x();
        cxyz[3 * s + 1] = p.y();
        cxyz[3 * s + 2] = p.z();
    }
    std::vector<int32_t> order(S);
    RBFX_CALL(ctx, rbfx_order_morton((int32_t)S, cxyz.data(), order.data()));
    int32_t* d_order = nullptr;
    RBFX_CALL(ctx, rbfx_dev_alloc(ctx, S * sizeof(int32_t), (void**)&d_order));
    RBFX_CALL(ctx, rbfx_h2d(ctx, d_order, order.data(), S * sizeof(int32_t)));
    rc = rbfx_ell_set_row_order(e, d_order);
    rbfx_dev_free(ctx, d_order);
    if (rc != RBFX_OK) die(ctx, "rbfx_ell_set_row_order", rc);
    return e;
}

rbfx_ell* RBFFD_HIP::getGPUWeights(DerType which) {
    updateWeightsOnGPU(true);
    int i = getDerTypeIndx(which);
    if (!ell_by_type[i]) ell_by_type[i] = buildEll(which, storage);
    return ell_by_type[i];
}

void RBFFD_HIP::ensureWork(unsigned int n) {
    if (work_n >= n) return;
    rbfx_dev_free(ctx, d_u);
    rbfx_dev_free(ctx, d_out);
    RBFX_CALL(ctx, rbfx_dev_alloc(ctx, (size_t)n * 8, &d_u));
    RBFX_CALL(ctx, rbfx_dev_alloc(ctx, (size_t)n * 8, &d_out));
    work_n = n;
}

void RBFFD_HIP::updateFunctionOnGPU(unsigned int, unsigned int nb_vals, double* u, bool) {
    ensureWork(grid_ref.getNodeListSize() > nb_vals ? grid_ref.getNodeListSize() : nb_vals);
    if (storage == RBFX_F64) {
        RBFX_CALL(ctx, rbfx_h2d(ctx, d_u, u, (size_t)nb_vals * 8));
    } else {
        std::vector<float> f(u, u + nb_vals);
        RBFX_CALL(ctx, rbfx_h2d(ctx, d_u, f.data(), (size_t)nb_vals * 4));
    }
}

void RBFFD_HIP::applyWeightsForDeriv(DerType which, std::vector<double>& u, std::vector<double>& deriv, bool isChangedU) {
    unsigned int nb_stencils = grid_ref.getStencilsSize();
    if (deriv.size() < nb_stencils) deriv.resize(nb_stencils);
    applyWeightsForDeriv(which, 0, nb_stencils, &u[0], &deriv[0], isChangedU);
}

// All rows are applied on the device (as RBFFD_CL does); rows [start, start+nb) are returned in
// deriv[start .. start+nb), the entries the reference's loop writes (rbffd.cpp:669-688).
void RBFFD_HIP::applyWeightsForDeriv(DerType which, unsigned int start_indx, unsigned int nb_stencils, double* u,
                                     double* deriv, bool isChangedU) {
    rbfx_ell* e = getGPUWeights(which);
    unsigned int N = grid_ref.getNodeListSize(), S = grid_ref.getStencilsSize();
    ensureWork(N > S ? N : S);
    if (isChangedU) updateFunctionOnGPU(0, N, u, true);
    int ft = storage == RBFX_F64 ? RBFX_F64 : RBFX_F32;
    int ac = storage == RBFX_F64 ? RBFX_F64 : accum;
    RBFX_CALL(ctx, rbfx_spmv(e, 0, d_u, d_out, ft, ac));
    if (start_indx + (size_t)nb_stencils > S) {
        std::cout << "[RBFFD_HIP] ERROR: applyWeightsForDeriv rows [" << start_indx << ", "
                  << start_indx + (size_t)nb_stencils << ") exceed the " << S << " stencils" << std::endl;
        std::exit(EXIT_FAILURE);
    }
    if (ft == RBFX_F64) {
        RBFX_CALL(ctx, rbfx_d2h(ctx, deriv + start_indx, (double*)d_out + start_indx, (size_t)nb_stencils * 8));
    } else {
        std::vector<float> f(nb_stencils);
        RBFX_CALL(ctx, rbfx_d2h(ctx, f.data(), (float*)d_out + start_indx, (size_t)nb_stencils * 4));
        for (unsigned i = 0; i < nb_stencils; i++) deriv[start_indx + i] = f[i];
    }
}

}  // namespace rbfx
