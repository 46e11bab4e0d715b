"""implicit_dpd_amd: MI355X-native MDPD engine (HIP kernels behind a C ABI)."""
