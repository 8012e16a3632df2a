"""schwarmax_amd: MI355X-native SchwarMAX model evaluation (HIP kernels via a C ABI).

Drop-ins for the reference's hot path:
    model.build_model, model.get_model_with_orbit   (model.py:70, 606)
    likelihood.model_likelihood                      (likelihood.py:96)
    nnls.solve_nnls_admm[_bootstrap]                 (nnls.py:6, 68)
"""
__all__ = ['model', 'likelihood', 'nnls', 'potentials', 'synth', 'utils']
