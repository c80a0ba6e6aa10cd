"""dbgenzpaper_amd: MI355X-native Genz quasi-Monte Carlo multivariate normal
probabilities (the qmc_pnorm! path of JakobPeterlin/DBGenzPaper).

The host API mirrors the reference's Julia API (QMC_opts, QMCData,
QMCDataSparse, QMCData_Distributions, qmc_pnorm!, cholesky_genz!,
cholesky_rowmax!, adapt_low_rank!, cholesky_classic!, SparseQMC); all numerical work runs in libdbgenz.so
(HIP, gfx950) behind the C ABI of include/dbgenz.h.
"""
from .qmc import (CholeskyGenz, QMC_opts, QMCData, QMCData_Distributions, QMCDataSparse,
                  QMCOpts, SparseQMC, adapt_low_rank, cholesky_classic, covariance, cholesky_genz, cholesky_rowmax, draw_shifts, qmc_pnorm)

__all__ = ["CholeskyGenz", "QMC_opts", "adapt_low_rank", "covariance", "QMCData", "QMCData_Distributions", "QMCDataSparse",
           "QMCOpts", "SparseQMC", "cholesky_classic", "cholesky_genz", "cholesky_rowmax", "draw_shifts",
           "qmc_pnorm"]
