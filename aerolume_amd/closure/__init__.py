"""Learned eddy-viscosity closure for the 3D solver: a small BF16 CNN trained on the GPU by
the package's own kernels on box-filtered fine-grid solves (``network``, ``data``, ``train``)."""
