"""The reference's own test-case lists and tolerance tables, restated as data.

Each case: name, kind, n (or nx, ny), batch, mode, and the reference's (atol, rtol) from its
compare tables.  Inputs are generated the way the reference's gen_data.py scripts do:
torch.manual_seed(seed) then torch.randn(batch, n, dtype=complex64) for complex input
(src/fft1_d/tests/fft1_d_data/gen_data.py:64-65), torch.randn(batch*n).float() for real input
(src/rfft1_d/tests/rfft1_d_data/gen_data.py:81-84), and for C2R a complex randn half-spectrum or
the Nyquist-only spectrum (src/irfft1_d/tests/irfft1_d_data/gen_data.py:51-58).  Goldens are
torch.fft on the CPU with the reference's conventions (ifft*n, irfft*n, ifft2*nx*ny).
"""

SEED = 12345  # src/common/host/test_helper.py:194

# src/fft1_d/tests/fft1_d_data/gen_data.py:25-56 and compare_data.py:19-47
C2C = [
    ("c2c_forward_n512", 512, 1, "forward", 1e-2, 1e-3),
    ("c2c_forward_n420", 420, 1, "forward", 1e-2, 1e-3),
    ("c2c_roundtrip_n512", 512, 1, "roundtrip", 0.512, 1e-2),
    ("dft_forward", 4, 1, "forward", 1e-3, 1e-3),
    ("dft_backward", 4, 1, "backward", 1e-3, 1e-3),
    ("dft_config", 256, 10, "forward", 1e-2, 1e-3),
    ("stride_forward", 256, 128, "vertical", 1e-2, 1e-3),
    ("stride_forward_n4096", 4096, 128, "vertical", 1e-2, 1e-3),
    ("stride_forward_bigstride", 256, 263168, "vertical", 1e-2, 1e-3),
    ("fft_n_forward", 32768, 1, "forward", 1e-2, 1e-3),
    ("fft_n_roundtrip", 32768, 1, "roundtrip", 3.2768, 1e-2),
    ("fft_n_config", 32768, 10, "forward", 1e-2, 1e-3),
    ("fft_b_n512", 512, 2, "forward", 1e-2, 1e-3),
    ("fft_b_n2048", 2048, 2, "forward", 1e-2, 1e-3),
    ("fft_b_n16384", 16384, 2, "forward", 1e-2, 1e-3),
    ("mix_fwd_n6", 6, 1, "forward", 1e-2, 1e-3),
    ("mix_fwd_n105", 105, 1, "forward", 1e-2, 1e-3),
    ("mix_fwd_n210", 210, 1, "forward", 1e-2, 1e-3),
    ("mix_bwd_n15", 15, 1, "backward", 1e-2, 1e-3),
    ("mix_prime_n11", 11, 1, "forward", 1e-2, 1e-3),
    ("mix_primepow_n121", 121, 1, "forward", 1e-2, 1e-3),
    ("mix_mixed_n143", 143, 1, "forward", 1e-2, 1e-3),
    ("mix_all4_n2310", 2310, 1, "forward", 1e-2, 1e-3),
    ("mix_vlarge_n46200", 46200, 1, "forward", 1e-2, 1e-3),
    ("mix_roundtrip_n210", 210, 1, "roundtrip", 2.1, 1e-2),
    ("mix_bwd_n210_b4", 210, 4, "backward", 1e-2, 1e-3),
    ("mix_fwd_n1536_b4", 1536, 4, "forward", 1e-2, 1e-3),
]

# src/rfft1_d/tests/rfft1_d_data/gen_data.py:25-78 and compare_data.py:19-70
R2C = [
    ("r2c_basic", 8, 1, 1e-3, 1e-3), ("r2c_n16", 16, 1, 1e-3, 1e-3), ("r2c_n1024", 1024, 1, 1e-2, 1e-3),
    ("r2c_dft_n7", 7, 1, 1e-3, 1e-3), ("r2c_dft_n15", 15, 1, 1e-3, 1e-3), ("r2c_dft_n255", 255, 1, 1e-2, 1e-3),
    ("r2c_n2048", 2048, 1, 1e-2, 1e-3), ("r2c_n4096", 4096, 1, 1e-2, 1e-3), ("r2c_n2100", 2100, 1, 1e-2, 1e-3),
    ("r2c_n999", 999, 1, 1e-2, 1e-3), ("r2c_n1023", 1023, 1, 1e-2, 1e-3), ("r2c_n5040", 5040, 1, 5e-2, 1e-2),
    ("r2c_n25200", 25200, 1, 2e-1, 1e-2), ("r2c_fft_n1536", 1536, 4, 1e-2, 1e-3),
    ("r2c_fft_n8192", 8192, 4, 5e-2, 1e-2), ("r2c_fft_n2048_b100", 2048, 100, 1e-2, 1e-3),
    ("r2c_fft1_1", 2, 3, 1e-3, 1e-3), ("r2c_fft1_2", 4, 16384, 1e-3, 1e-3), ("r2c_fft1_3", 8, 1, 1e-3, 1e-3),
    ("r2c_fft1_4", 128, 16384, 1e-3, 1e-3), ("r2c_fft2_1", 256, 3, 1e-3, 1e-3),
    ("r2c_fft2_2", 512, 16384, 1e-3, 1e-3), ("r2c_fft2_3", 4096, 1, 1e-3, 1e-3),
    ("r2c_fft2_4", 16384, 16384, 1e-3, 1e-3), ("r2c_fft3_1_1", 32768, 3, 1e-3, 1e-3),
    ("r2c_fft3_1_2", 32768, 4096, 1e-3, 1e-3), ("r2c_fft3_1_3", 65536, 1, 1e-3, 1e-3),
    ("r2c_fft3_1_4", 65536, 4096, 1e-3, 1e-3), ("r2c_fft3_2_1", 131072, 3, 1e-3, 1e-3),
    ("r2c_fft3_2_2", 262144, 2048, 1e-3, 1e-3), ("r2c_fft3_2_3", 1048576, 1, 1e-2, 1e-2),
    ("r2c_fft3_2_4", 2097152, 64, 1e-2, 1e-2), ("r2c_fft4_1", 4194304, 3, 1e-1, 1e-2),
    ("r2c_fft4_2", 8388608, 16, 1e-1, 1e-2), ("r2c_fft4_3", 16777216, 1, 1e-1, 1e-2),
    ("r2c_fft4_4", 33554432, 8, 1e-1, 1e-2), ("r2c_fft5_1", 67108864, 4, 1e-1, 1e-2),
    ("r2c_fft5_2", 134217728, 2, 1e-0, 1e-2), ("r2c_fft3_2_direct_b200", 16384, 200, 1e-3, 1e-3),
]

# src/irfft1_d/tests/irfft1_d_data/gen_data.py:25-44 and compare_data.py:19-37
C2R = [
    ("c2r_basic", 8, 1, False, 1e-3, 1e-3), ("c2r_n16", 16, 1, False, 1e-3, 1e-3),
    ("c2r_dft_n7", 7, 1, False, 1e-3, 1e-3), ("c2r_dft_n15", 15, 1, False, 1e-3, 1e-3),
    ("c2r_dft_n255", 255, 1, False, 1e-2, 1e-3), ("c2r_n2100", 2100, 1, False, 1.0, 1e-2),
    ("c2r_n5040", 5040, 1, False, 5.0, 1e-2), ("c2r_fft_n1536", 1536, 4, False, 5e-2, 1e-2),
    ("c2r_fft_n8192", 8192, 4, False, 5e-2, 1e-2), ("c2r_fft_n8192_nyquist", 8192, 1, True, 5e-2, 1e-2),
    ("c2r_fft_n2048_b100", 2048, 100, False, 5e-2, 1e-2), ("c2r_fft_n12288", 12288, 1, False, 1e-3, 5e-4),
    ("c2r_fft_n8190", 8190, 1, False, 1e-3, 5e-4), ("c2r_fft_odd_n1025", 1025, 4, False, 5e-2, 1e-2),
]

# src/fft2_d/tests/fft2_d_data/gen_data.py:25-44 and compare_data.py:19-35
FFT2 = [
    ("dd_32x32_forward", 32, 32, 1, True, 1e-2, 1e-3), ("dd_32x64_forward", 32, 64, 1, True, 1e-2, 1e-3),
    ("dd_64x64_forward", 64, 64, 1, True, 1e-2, 1e-3), ("dd_64x128_forward", 64, 128, 1, True, 1e-2, 1e-3),
    ("dd_128x128_forward", 128, 128, 1, True, 1e-2, 1e-3), ("dd_32x32_backward", 32, 32, 1, False, 1.0, 1e-2),
    ("dd_64x64_backward", 64, 64, 1, False, 1.0, 1e-2), ("dd_128x128_backward", 128, 128, 1, False, 1.0, 1e-2),
    ("dd_128x32_forward_b2", 128, 32, 2, True, 1e-2, 1e-3), ("dd_64x64_forward_b2", 64, 64, 2, True, 1e-2, 1e-3),
    ("dd_64x32_forward_b4", 64, 32, 4, True, 1e-2, 1e-3), ("dd_32x128_forward_b2", 32, 128, 2, True, 1e-2, 1e-3),
    ("dd_32x64_forward_b4", 32, 64, 4, True, 1e-2, 1e-3), ("dd_32x32_forward_b8", 32, 32, 8, True, 1e-2, 1e-3),
    ("dd_64x64_backward_b2", 64, 64, 2, False, 1.0, 1e-2),
]
