"""Coefficient lists (module c_data, DBSR_MAT3/MOD_c_data.f90) -- see capi."""
