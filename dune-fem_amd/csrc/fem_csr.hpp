// Assembled Jacobian of a Galerkin operator in CSR form on the device.
#pragma once

#include "fem_op.hpp"

namespace femamd {

struct CSR {
    Context *ctx = nullptr;
    Operator *op = nullptr;
    int64_t N = 0, nnz = 0;
    DevBuf rowptr;   // int64 [N + 1]
    DevBuf col;      // int32 [nnz], increasing within a row
    DevBuf val;      // FP64 [nnz]
    void assemble(Operator *o);
    void apply(const Vec &x, Vec &y) const;   // y = A x (FP64 or FP32 vectors, FP64 accumulation)
    void diagonal(Vec &d) const;              // d_i = A_ii
    void download(int64_t *rowptr, int *col, double *val) const;
};

}  // namespace femamd

struct fem_csr { femamd::CSR m; femamd::Context *held = nullptr; };
