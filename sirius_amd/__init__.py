"""MI355X-native decoder of DuckDB storage segments (BITPACKING, RLE, ALP,
ALPRD, FSST) behind the C ABI in include/sdr/decode.h."""
__version__ = "0.1.0"
