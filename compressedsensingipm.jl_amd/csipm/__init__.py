"""csipm: an MI355X (gfx950) implementation of the matrix-free interior-point
hot path of CompressedSensingIPM.jl -- the condensed FFT-KKT operator apply,
the beta <-> DFT maps and the CG vector work -- behind the C ABI of
libcsipm.so, with Python mirrors of the reference's FFTParameters /
FFTOperator / FFTNLPModel / FFTKKTSystem objects, a batch engine for uneven
instances and a continuous-batching serving loop (csipm.batch)."""
