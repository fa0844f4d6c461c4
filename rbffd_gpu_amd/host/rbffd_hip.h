// RBFFD interface (src/rbffd/rbffd.h:29-547) and its MI355X implementation RBFFD_HIP, which
// takes the place of RBFFD_CL / RBFFD_VCL: weights are computed on the GPU by one batched
// call (rbfx_weights), kept resident, and applied there (rbfx_spmv / rbfx_spmm).
//
// Kept from the reference: DerType bit values and their ascending-bit column order, the
// constructor signature, the epsilon / hyperviscosity setters and their formulas, both
// applyWeightsForDeriv overloads, the jagged host weight storage weights[type][stencil]
// (n entries per stencil), getWeights / getStencilWeights, and exit-on-error behaviour.
// Differences: computeAllWeightsForAllStencils is virtual here (the reference's is not, so a
// subclass cannot override it through an RBFFD&), and the base class is abstract — the CPU
// implementation of the reference is the test-side restatement, not part of this library.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "grid.h"
#include "rbfx.h"

namespace rbfx {

class RBFFD {
  public:
    enum DerType {
        X = 0x1, Y = 0x2, Z = 0x4, LAPL = 0x8, R = 0x10, HV = 0x20, LAMBDA = 0x40, THETA = 0x80,
        LSFC = 0x100, XSFC = 0x200, YSFC = 0x400, ZSFC = 0x800, ALT_XSFC = 0x1000, ALT_YSFC = 0x2000,
        ALT_ZSFC = 0x4000, INTERP = 0x8000
    };
    enum DerTypeIndx {
        X_i = 0, Y_i, Z_i, LAPL_i, R_i, HV_i, LAMBDA_i, THETA_i, LSFC_i, XSFC_i, YSFC_i, ZSFC_i, ALT_XSFC_i,
        ALT_YSFC_i, ALT_ZSFC_i, INTERP_i, NUM_DERIVATIVE_TYPES
    };
    typedef unsigned int DerTypes;
    enum WeightType { Direct, ContourSVD };

    RBFFD(DerTypes typesToCompute, Grid* grid, int dim_num, int rank = 0);
    virtual ~RBFFD();

    virtual void computeAllWeightsForAllStencils() = 0;
    virtual void applyWeightsForDeriv(DerType which, std::vector<double>& u, std::vector<double>& deriv,
                                      bool isChangedU = true) = 0;
    virtual void applyWeightsForDeriv(DerType which, unsigned int start_indx, unsigned int nb_stencils, double* u,
                                      double* deriv, bool isChangedU = true) = 0;

    // rbffd.h:213-245
    void setUseHyperviscosity(int tf);
    int getUseHyperviscosity() const { return useHyperviscosity; }
    double getHVScalar() const;  // -hv_gamma * N^-hv_k
    double setHVScalars(int k, double gamma_c) {
        hv_gamma = gamma_c;
        hv_k = k;
        return hv_gamma;
    }
    void setWeightType(WeightType t);
    // rbffd.h:285-375
    void setEpsilonByParameters(double c1, double c2);
    void setEpsilonByStencilSize();
    void setEpsilon(double eps);
    void setEpsilon(std::vector<double>& lst_of_epsilon);
    const std::vector<double>& getEpsilon() const { return var_epsilon; }
    // Polynomial augmentation degree (the reference fixes the single constant monomial,
    // rbffd.cpp:1 ONE_MONOMIAL; degrees 1..3 are this build's addition).
    void setPolynomialDegree(int deg) { poly_deg = deg; }
    int getPolynomialDegree() const { return poly_deg; }

    std::vector<double*>& getWeights(DerType choice) { return weights[getDerTypeIndx(choice)]; }
    double*& getStencilWeights(DerType choice, int st_indx) { return weights[getDerTypeIndx(choice)][st_indx]; }
    DerTypes getComputedTypes() const { return computedTypes; }
    void appendDerTypes(DerTypes t) { computedTypes |= t; }
    void removeDerType(DerTypes t) { computedTypes &= ~t; }
    int getDimNum() const { return dim_num; }
    Grid& getGrid() { return grid_ref; }

    // MatrixMarket weight files, ascii "%d %d %24.16le" 1-based (rbffd.cpp:1237-1300) or the
    // binary form: banner + size line, then (int row, int col, double w) triples (1302-1375).
    void setAsciiWeights(int ascii) { asciiWeights = ascii; }
    void writeToFile(DerType which, const std::string& filename);
    int loadFromFile(DerType which, const std::string& filename);

    static DerType getDerType(int i) { return (DerType)(0x1 << i); }
    static DerTypeIndx getDerTypeIndx(DerType dt) {
        int it = dt, i = -1;
        while (it) {
            it >>= 1;
            i++;
        }
        return (DerTypeIndx)i;
    }

  protected:
    DerTypes computedTypes;
    Grid& grid_ref;
    int dim_num;
    int rank;
    std::vector<double*> weights[NUM_DERIVATIVE_TYPES];
    std::vector<double> var_epsilon;
    double hv_k = 2;
    double hv_gamma = 8e-4;
    int useHyperviscosity = 0;
    int poly_deg = 0;
    int asciiWeights = 0;
    bool weightsModified = false;
    void storeHostWeights(const std::vector<double>& w_packed, const std::vector<int32_t>& st_off, DerTypes types);
};

class RBFFD_HIP : public RBFFD {
  public:
    RBFFD_HIP(DerTypes typesToCompute, Grid* grid, int dim_num, int rank = 0);
    virtual ~RBFFD_HIP();

    // All stencils in one batched GPU call (weights stay on the device; host copies are filled
    // for getWeights / file output).
    void computeAllWeightsForAllStencils() override;
    void applyWeightsForDeriv(DerType which, std::vector<double>& u, std::vector<double>& deriv,
                              bool isChangedU = true) override;
    void applyWeightsForDeriv(DerType which, unsigned int start_indx, unsigned int nb_stencils, double* u,
                              double* deriv, bool isChangedU = true) override;

    // RBFFD_CL extras (rbffd_cl.h:76-145)
    void updateWeightsOnGPU(bool forceFinish = true);
    void updateFunctionOnGPU(unsigned int start_indx, unsigned int nb_vals, double* u, bool forceFinish = true);
    rbfx_ell* getGPUWeights(DerType which);          // single-operator set of one type
    rbfx_ell* getGPUOperatorSet(DerTypes types);      // fused set (ascending bit order)
    rbfx_ell* getGPUOperatorSet(DerTypes types, int storage_);  // the same in a given storage
    const int32_t* getGPUStencils() const { return d_idx; }
    const double* getGPUNodes() const { return d_xyz; }
    unsigned int getStencilPaddedSize() const { return width; }
    bool areGPUKernelsDouble() const { return storage == RBFX_F64; }
    // Precision of the device operators: RBFX_F64 (default) or RBFX_F32 storage; accumulation
    // RBFX_F64 or RBFX_F32 (the reference's float mode is storage F32 + accumulation F32).
    void setGPUPrecision(int storage_, int accum_);
    void setSolveMode(int mode) { solve_mode = mode; }
    rbfx_ctx* getContext() { return ctx; }
    // Per-stencil diagnostics of the last weight solve: {backward error, growth, iterations, path}.
    const std::vector<double>& getSolveDiagnostics() const { return diag; }

  private:
    void uploadGeometry();
    void ensureWork(unsigned int n);
    rbfx_ell* buildEll(DerTypes types, int storage_);  // padded weights of the types -> device operator set
    rbfx_ctx* ctx = nullptr;
    int storage = RBFX_F64, accum = RBFX_F64, solve_mode = RBFX_SOLVE_F64_LU;
    unsigned int width = 0;
    double* d_xyz = nullptr;
    int32_t* d_idx = nullptr;      // padded n_st x width (pad = centre)
    int32_t* d_off = nullptr;      // CSR offsets of the (possibly ragged) stencils
    int32_t* d_csr = nullptr;      // CSR indices
    double* d_w = nullptr;         // T x nnz packed weights from the solve
    double* d_wpad = nullptr;      // one padded weight array (staging for ELL builds)
    std::vector<int32_t> st_off;
    std::vector<rbfx_ell*> ell_by_type = std::vector<rbfx_ell*>(NUM_DERIVATIVE_TYPES, nullptr);
    std::vector<std::pair<uint64_t, rbfx_ell*>> fused;  // key: types | storage << 32
    void* d_u = nullptr;
    void* d_out = nullptr;
    unsigned int work_n = 0;
    std::vector<double> diag;
};

}  // namespace rbfx
