// Launch table of the precompiled kernel variants.  One table per (velocity set,
// storage, collision); filled by the per-velocity-set translation units.
#pragma once
#include "lbm_device.hpp"

namespace fx {

struct LaunchShape {
	dim3 grid, block;
	size_t lds = 0;       // dynamic LDS reserved per block (occupancy limit), see occupancy_lds()
	uint32_t xchunks = 0; // 1-D XCD-mapped grid with this many x chunks per row (see block_coords)
	bool lin = false;     // cell-linear grid (cell_coords in lbm_device.hpp), see linear_shape
	bool pairs = false;   // two cells per lane (k_stream_collide_pair), see pair_shape
	bool off32 = false;   // row grid with 32-bit slot offsets (O32 of load_raw/store_f)
	bool buf = false;     // row grid through buffer instructions (BUF of load_raw/store_f)
	bool alternate = false; // stream_collide takes its blocks in reverse order on odd steps
};
// A launch holds at most 2^32-1 work-items along each grid dimension (grid size times
// block size); lattices of more than ~4.29 G cells (a D3Q19 FP16 lattice of up to ~5.2 G
// cells fits in 288 GB) must not take a 1-D grid over all cells.
constexpr uint64_t MAX_ITEMS_PER_DIMENSION = 0xFFFFFFFFull;
inline bool grid_fits(uint64_t blocks, uint32_t block_threads) { return blocks * block_threads <= MAX_ITEMS_PER_DIMENSION; }
// Every row of every DDF slot starts on a 128-byte line when Nx times the element size is
// a multiple of 128 bytes (the slot stride N times the element size then is one too, and
// device buffers are allocated 256-byte aligned).
inline bool rows_on_lines(uint32_t Nx, uint32_t elem_bytes) { return (uint64_t)Nx * elem_bytes % 128u == 0u; }
// stream_collide and the adjoint kernels on lattices whose rows miss the lines: blocks of
// 256 consecutive cells of the whole lattice.  Measured (plain D3Q19, against blocks over
// rows with cached accesses): 200^3 FP32 +21 %, FP16S +24 %; 300^3 +16 %, +23 %;
// 1000x250x250 +12 %, +3 %; 352x352x256 FP16S +19 %.  xcd pads the grid to whole groups
// of 64 blocks for the XCD-aware segment order (linear_segment).
inline uint64_t linear_blocks(uint64_t N, bool xcd) {
	const uint64_t blocks = (N + 255u) / 256u;
	return xcd ? (blocks + 63u) / 64u * 64u : blocks;
}
inline LaunchShape linear_shape(uint64_t N, bool xcd) {
	LaunchShape s;
	const uint64_t blocks = linear_blocks(N, xcd);
	s.block = dim3(256, 1, 1);
	s.grid = dim3((unsigned)blocks, 1, 1);
	s.lin = true;
	return s;
}
// Block of bx consecutive x cells of one row (bx = 64, 128 or 256); grid over (x chunk,
// y, z).  bx is the widest block that leaves less than one idle wave per row (64 when none
// does).
inline LaunchShape lattice_shape(uint32_t Nx, uint32_t Ny, uint32_t Nz) {
	uint32_t bx = 64u;
	for(uint32_t b = 256u; b > 64u; b >>= 1)
		if((Nx + b - 1u) / b * b - Nx < 64u) { bx = b; break; }
	LaunchShape s;
	s.block = dim3(bx, 1, 1);
	s.grid = dim3((Nx + bx - 1u) / bx, Ny, Nz);
	return s;
}

// stream_collide: with xcd_map, rows that take more than one block use the 1-D XCD-mapped
// grid, which keeps all chunks of a row on one XCD (block_coords in lbm_device.hpp).
inline LaunchShape stream_shape(uint32_t Nx, uint32_t Ny, uint32_t Nz, bool xcd_map) {
	LaunchShape s = lattice_shape(Nx, Ny, Nz);
	const uint32_t chunks = s.grid.x;
	const uint64_t groups = ((uint64_t)Ny * Nz + 7u) / 8u;
	// otherwise the 3-D row grid: (chunks, Ny, Nz) stays within every dimension's limit
	if(xcd_map && chunks > 1u && grid_fits(groups * 8u * chunks, s.block.x)) {
		s.grid = dim3((unsigned)(groups * 8u * chunks), 1, 1);
		s.xchunks = chunks;
	}
	return s;
}

// k_stream_collide_pair runs on rows whose Nx/2 lanes fill whole waves (Nx a multiple of
// 128, so the 16-bit rows also start on 128-byte lines).  Rows of half a wave left a
// quarter of the lanes idle: 192x64x64 FP16S ran 12 % below one cell per lane.
inline bool pair_rows(uint32_t Nx) { return Nx % 128u == 0u; }
// Its grid: the row grid over Nx/2 lanes per row, XCD-mapped when a row takes several blocks.
inline LaunchShape pair_shape(uint32_t Nx, uint32_t Ny, uint32_t Nz) {
	LaunchShape s = stream_shape(Nx / 2u, Ny, Nz, true);
	s.pairs = true;
	return s;
}

// Dynamic LDS per block that caps residency at `waves_per_simd` waves per SIMD (4 SIMDs
// per CU), for a kernel that uses no LDS itself.  0 when the cap would not bind exactly.
// A block's LDS is allocated in granules of 1/128 of the CU's LDS (1280 B of 160 KiB on
// MI355X; measured: 53760 B per 256-thread block leaves 3 blocks per CU resident, 54272 B
// only 2), so the size is a whole number of granules, at most the per-block limit (2 blocks
// of 256 threads: 65280 B).
inline size_t occupancy_lds(uint32_t waves_per_simd, uint32_t block_threads, size_t lds_per_cu, size_t lds_per_block_max) {
	const uint32_t waves_per_block = (block_threads + 63u) / 64u;
	const uint32_t blocks = (waves_per_simd * 4u) / waves_per_block;
	const size_t granule = lds_per_cu / 128u;
	if(blocks == 0u || granule == 0u) return 0;
	const size_t share = lds_per_cu / blocks;
	const size_t lds = (share < lds_per_block_max ? share : lds_per_block_max) / granule * granule;
	if(lds == 0u || lds > lds_per_block_max || lds_per_cu / lds != blocks) return 0;
	return lds;
}

// Grid of the adjoint kernels (record pass and adjoint step): XCD-mapped, full occupancy.
inline LaunchShape adjoint_shape(uint32_t Nx, uint32_t Ny, uint32_t Nz) { return stream_shape(Nx, Ny, Nz, true); }

typedef void (*lattice_fn)(const LbmArgs&, const LaunchShape&, hipStream_t);
typedef void (*mass_fn)(const LbmArgs&, const LaunchShape&, hipStream_t, double*);
typedef void (*record_fn)(const LbmArgs&, const LaunchShape&, hipStream_t, float*);
typedef void (*adjoint_fn)(const AdjointArgs&, const LaunchShape&, hipStream_t);
typedef void (*objective_fn)(const LbmArgs&, uint32_t, hipStream_t, double*);
typedef void (*transfer_fn)(void* fi, void* buffer_p, void* buffer_m, uint32_t Nx, uint32_t Ny, uint32_t Nz, uint32_t axis, bool insert,
                            uint32_t odd, hipStream_t);

struct LbmKernels {
	lattice_fn initialize;
	lattice_fn stream_collide;
	lattice_fn update_fields;
	lattice_fn update_force_field;
	mass_fn mass_partial;
	uint32_t stream_collide_waves_cap;
	record_fn stream_collide_record;
	adjoint_fn adjoint_step;
	objective_fn objective_partial; // grid of ceil(Ny*Nz/256) blocks, one partial each
	bool stream_collide_xcd_map;    // 1-D XCD-mapped grid for rows of several blocks
	transfer_fn transfer_fi;        // halo transfer of the DDFs (k_transfer_fi)
	lattice_fn update_moving_boundaries;
	bool stream_collide_linear_large; // cell-linear grid for DDF arrays above LARGE_DDF_BYTES
	bool stream_collide_linear_xcd;   // XCD-aware segment order on the cell-linear grid
	bool stream_collide_pairs;        // k_stream_collide_pair on rows that start on lines
	bool stream_collide_offsets32;    // 32-bit slot offsets on the row grid (slots below 4 GiB)
	bool stream_collide_buffer;       // buffer instructions on the row grid (DDFs below 4 GiB)
	bool stream_collide_pair_buffer;  // the same for k_stream_collide_pair
	bool stream_collide_linear_buffer; // the same on the cell-linear grid
};

constexpr uint32_t EXT_ALL = EXT_VOLUME_FORCE | EXT_FORCE_FIELD | EXT_EQUILIBRIUM_BOUNDARIES | EXT_UPDATE_FIELDS | EXT_POROSITY |
                             EXT_PRECISE_DENSITY | EXT_MOVING_BOUNDARIES | EXT_SUBGRID;

// Measured launch tuning (see KernelSet), MI355X, 256^3 to 1536x768x768:
// - FP32 kernels (non-temporal DDF access) stream fastest with fewer concurrent DDF
//   streams: plain D3Q19/D3Q27 at 3 waves/SIMD (+11 % / +3 % over full occupancy),
//   D3Q19 with EQUILIBRIUM_BOUNDARIES (+FORCE_FIELD) at 4 (+2-4 %), D3Q19 with
//   VOLUME_FORCE (+POROSITY) at 3 (+4 %); the generic all-extension kernel is not capped;
// - the 2-byte formats (non-temporal too) need full occupancy;
// - on rows of several blocks the XCD-mapped grid gains 9-24 % for FP16S, 6-10 % for FP32
//   (plain and porous, 384x192x192 to 768x384x192) and 5-10 % for FP16C (512x256x256,
//   plain D3Q19 and the D3Q27 TRT EQUILIBRIUM_BOUNDARIES + FORCE_FIELD kernel);
// - one block per row (up to 768 threads) measured no better than chunks of 128/256;
// - two or four cells per lane, each with accesses of its own, measured no better than
//   one for any variant once all loads are issued ahead of the arithmetic; two cells that
//   share 32-bit words are tune_pairs' kernel.
constexpr uint32_t tune_waves_cap(int Q, uint32_t ST, uint32_t EXT) {
	if(ST != ST_FP32) return 0u;
	if(EXT == 0u) return 3u;
	if(Q != 19) return 0u;
	if(EXT == EXT_EQUILIBRIUM_BOUNDARIES || EXT == (EXT_EQUILIBRIUM_BOUNDARIES | EXT_FORCE_FIELD)) return 4u;
	if(EXT == EXT_VOLUME_FORCE || EXT == (EXT_VOLUME_FORCE | EXT_POROSITY | EXT_PRECISE_DENSITY)) return 3u;
	return 0u;
}
constexpr bool tune_xcd_map(int Q, uint32_t ST, uint32_t EXT) {
	return true;
}
// 32-bit slot offsets on the row grid (load_raw's O32: one base per slot in scalar registers
// and one 32-bit offset per lane instead of 64-bit per-lane addresses): the D3Q27 16-bit
// kernels spend about 45 of their ~850 VALU instructions per wave on 64-bit address
// arithmetic, and the C3-type kernel (TRT FP16C, EQUILIBRIUM_BOUNDARIES + FORCE_FIELD) keeps
// the VALU 99 % busy.  Measured at 256^3 against 64-bit addressing, same process, bit-exact:
// SRT none +4.6 % (FP16S) / +2.3 % (FP16C), EQUILIBRIUM_BOUNDARIES +3.7 / +2.2,
// VOLUME_FORCE +2.9 / -0.7, + FORCE_FIELD +2.9 / +1.6, porous +1.3 / +2.1; TRT
// EQUILIBRIUM_BOUNDARIES +2.6 / +1.3, + FORCE_FIELD +2.7 / +0.7 (+2.5 % for FP16C at
// 512x256x256), porous +0.2 / +2.4; TRT none -3.3 / -4.0 and VOLUME_FORCE -3.6 / -4.5 and the
// all-extension kernels -4.0 to +0.5 % keep 64-bit addresses.  D3Q19 16-bit kernels that do
// not take two cells per lane: TRT none +5.2 / +1.4 %, EQUILIBRIUM_BOUNDARIES +4.3 / +0.5,
// + FORCE_FIELD +3.4 / -0.2, porous +1.6 / +4.5 take 32-bit offsets; TRT VOLUME_FORCE
// (-1.0 / -4.4), SRT VOLUME_FORCE and porous (-16 to +2 %) and the all-extension kernels
// (-15 to -5 %) keep their addressing (uniform row bases where uniform_rows says so).
constexpr bool tune_offsets32(int Q, uint32_t ST, bool TRT, uint32_t EXT) {
	if(ST == ST_FP32 || EXT == EXT_ALL) return false;
	if(Q == 27) return !TRT || (EXT != 0u && EXT != EXT_VOLUME_FORCE);
	return TRT && EXT != EXT_VOLUME_FORCE;
}
// Two cells per lane (k_stream_collide_pair) for the D3Q19 16-bit kernels where it measured
// faster than one cell per lane (256^3, same process, bit-exact; FP16S / FP16C): SRT none,
// EQUILIBRIUM_BOUNDARIES and + FORCE_FIELD +8 to +13 %; TRT none +1.9 / +4.1 %,
// EQUILIBRIUM_BOUNDARIES +3.6 / +0.3, + FORCE_FIELD +3.9 / +0.3, porous +1.5 / +7.3; SRT
// VOLUME_FORCE +1.6 / -0.7 and SRT porous -0.3 / +3.7 (each for the storage it gains on).
// TRT VOLUME_FORCE (-1.3 / +0.1), the all-extension kernels (SRT -7 to -8 %; and the pair
// kernel has no MOVING_BOUNDARIES) and D3Q27 (133-156 VGPRs, -13 to +0.3 %) keep one cell per lane.
constexpr bool tune_pairs(int Q, uint32_t ST, bool TRT, uint32_t EXT) {
	constexpr uint32_t POROUS = EXT_VOLUME_FORCE | EXT_POROSITY | EXT_PRECISE_DENSITY;
	if(Q != 19 || ST == ST_FP32 || (EXT & (EXT_MOVING_BOUNDARIES | EXT_SUBGRID)) != 0u) return false;
	if(EXT == EXT_VOLUME_FORCE) return !TRT && ST == ST_FP16S;
	if(EXT == POROUS) return TRT || ST == ST_FP16C;
	return EXT == 0u || EXT == EXT_EQUILIBRIUM_BOUNDARIES || EXT == (EXT_EQUILIBRIUM_BOUNDARIES | EXT_FORCE_FIELD);
}
// Buffer instructions on the row grid (load_raw's BUF: one buffer resource, scalar slot and
// row offsets, three lane offsets) for the 16-bit kernels that keep one cell per lane: their
// vector registers fall from 59-100 to 43-85 (the C3-type D3Q27 TRT FP16C kernel with
// EQUILIBRIUM_BOUNDARIES + FORCE_FIELD 96 -> 48) with no 64-bit addresses held from the loads
// to the stores.  Measured at 256^3 against the kernel selected before (32-bit offsets where
// tune_offsets32 says so), same process, bit-exact (FP16S / FP16C): D3Q27 SRT
// EQUILIBRIUM_BOUNDARIES +0.9 / +3.0 %, + FORCE_FIELD +1.0 / +2.2, porous +11.4 / +15.1,
// all extensions +22.7 / +20.6, VOLUME_FORCE -1.1 / +4.8; D3Q27 TRT none +1.2 / +2.3,
// EQUILIBRIUM_BOUNDARIES +0.4 / +3.4, VOLUME_FORCE +3.6 / +3.4, + FORCE_FIELD +0.8 / +2.4
// (+3.4 % for FP16C at 512x256x256), porous +10.1 / +3.7, all +15.6 / +16.9; D3Q19 TRT
// VOLUME_FORCE +6.2 / +3.9, TRT all +2.8 / +4.6, SRT porous (FP16S) +1.1, SRT VOLUME_FORCE
// (FP16C) +2.4.  Not taken: D3Q27 SRT none (-3.3 / -3.7), D3Q27 SRT VOLUME_FORCE FP16S,
// D3Q19 SRT all extensions (-2.2 / -1.5), every FP32 kernel (-9 to +1 %: they wait on HBM),
// and the D3Q19 kernels of tune_pairs (two cells per lane; on rows of half a wave they keep
// the one-cell form they had).
constexpr bool tune_buffer(int Q, uint32_t ST, bool TRT, uint32_t EXT) {
	if(ST == ST_FP32) return false;
	if(Q == 27) return TRT || (EXT != 0u && !(EXT == EXT_VOLUME_FORCE && ST == ST_FP16S));
	return !tune_pairs(Q, ST, TRT, EXT) && (TRT || EXT != EXT_ALL);
}
// The two-cells-per-lane kernel through buffer instructions (its BUF): its stores no longer
// rebuild 64-bit addresses after the collision.  Measured against global addressing, same
// process, bit-exact (256^3, then 1024x128x96; FP16S / FP16C): SRT none -1.4 / +0.5 and
// -0.6 / +3.4 %, EQUILIBRIUM_BOUNDARIES (+ FORCE_FIELD) -1.0 / +2.4 and +0.8 / +3.5, SRT
// VOLUME_FORCE (FP16S) +5.6 and +4.1, SRT porous (FP16C) +1.0 and +2.7, TRT none +2.7 / +5.9
// and +4.6 / +6.3, TRT EQUILIBRIUM_BOUNDARIES (+ FORCE_FIELD) +3.9 to +9.3, TRT porous +2.6 to
// +3.0.  The FP16S SRT kernels without forcing keep global addressing.
constexpr bool tune_pair_buffer(int Q, uint32_t ST, bool TRT, uint32_t EXT) {
	const bool unforced = EXT == 0u || EXT == EXT_EQUILIBRIUM_BOUNDARIES || EXT == (EXT_EQUILIBRIUM_BOUNDARIES | EXT_FORCE_FIELD);
	return tune_pairs(Q, ST, TRT, EXT) && !(ST == ST_FP16S && !TRT && unforced);
}
// The cell-linear grid through buffer instructions (load_raw's BUF with SC: a scalar slot
// offset plus per-lane cell offsets).  Measured at 200^3 against that grid's 32-bit slot
// offsets, same process, bit-exact: every 16-bit kernel measured (D3Q19 and D3Q27, SRT and
// TRT, none / EQUILIBRIUM_BOUNDARIES / VOLUME_FORCE / + FORCE_FIELD / porous / all
// extensions) +2 to +30 % except D3Q19 FP16C TRT VOLUME_FORCE (-0.5 %); at 300^3 plain D3Q19
// +1.1 % (FP16S) / +3.8 % (FP16C), D3Q19 FP16S TRT EQUILIBRIUM_BOUNDARIES + FORCE_FIELD
// +3.9 %, the C3-type D3Q27 kernel +12.9 %.  FP32 measured -4.9 to +2.2 % and keeps global
// addressing.
constexpr bool tune_linear_buffer(int Q, uint32_t ST, bool TRT, uint32_t EXT) { return ST != ST_FP32; }
// Block order of stream_collide: forward on even steps, reversed on odd ones, so that the
// rows a step wrote last are the first the next step reads (each step updates every cell
// independently, so the order does not change a result).  Measured against the forward
// order every step, same process, bit-exact: 512x256x256 FP32 +4.0 %, FP16S +7.2 %, FP16C
// +6.0 %; the C3-type D3Q27 TRT FP16C kernel +1.3 % (512x256x256) and +0.8 % (256^3);
// 300^3 (cell-linear) FP16S +2.2 %, FP32 +1.7 %; 384x192x192 FP32 +0.6 %; 512^3 FP16S
// +2.6 %; 1024x512x512 FP16S +1.7 %, FP32 +0.4 %; 1536x768x768 FP16S +0.6 %; at 256^3 D3Q19
// -0.2 to +0.2 %, below it 128^3 FP16S -2.4 % and 256x256x128 FP16S -0.6 %.  Taken from
// 768 MiB of DDFs on.
constexpr uint64_t ALTERNATE_ORDER_BYTES = 768ull << 20;
// Above 4 GiB of DDFs the FP32 kernels stream faster on the cell-linear grid than on the
// row grid, also where the rows start on lines (same residency cap).  Plain D3Q19: 384^3
// +10 %, 256x512x512 +2 %, 512^3 +12 to +16 %, 768x512x512 +2.5 %, 1024x512x512 +11 %;
// the D3Q19 EQUILIBRIUM_BOUNDARIES, VOLUME_FORCE, +FORCE_FIELD and porous kernels +12 to
// +14 % at 512^3 and +4 to +5 % at 768x512x512; plain D3Q27 +0.4 to +1 %.  Below, the row
// grid is ahead (256^3 +10 %, 512x256x256 +2 %).  The 16-bit formats keep the row grid at
// every size (+1 to +7 %).
constexpr uint64_t LARGE_DDF_BYTES = 4ull << 30;
constexpr bool tune_linear_large(int Q, uint32_t ST, uint32_t EXT) { return ST == ST_FP32; }
// On the cell-linear grid the XCD-aware segment order gains 3-7 % for FP16S (200^3 to
// 1000x250x250, full occupancy) and -2 to +1 % for FP32 at 3 waves/SIMD.
constexpr bool tune_linear_xcd(int Q, uint32_t ST, uint32_t EXT) { return ST != ST_FP32; }
// stream_collide (and the adjoint kernels) on the cell-linear grid: rows that miss the
// 128-byte lines, or a large FP32 lattice.  The 16-bit kernels address a slot there with
// 32-bit offsets (slot_offsets32), so a 16-bit slot of 4 GiB or more (2^31 cells) keeps
// the row grid.
inline bool use_linear_grid(const LbmKernels& k, uint32_t Nx, uint32_t elem_bytes, uint64_t N, uint32_t q) {
	if(elem_bytes != 4u && N * elem_bytes > 0xFFFFFFFFull) return false;
	if(!grid_fits(linear_blocks(N, k.stream_collide_linear_xcd), 256u)) return false;
	return !rows_on_lines(Nx, elem_bytes) || (k.stream_collide_linear_large && N * q * elem_bytes > LARGE_DDF_BYTES);
}

// Kernel set of one (velocity set, storage, collision) variant; EXT = 0 instantiates the
// extension-free kernels, EXT_ALL the ones that read the runtime extension mask.
template<int Q, uint32_t ST, bool TRT, uint32_t EXT> struct KernelSet {
	static void initialize(const LbmArgs& a, const LaunchShape& s, hipStream_t st) { k_initialize<Q, ST><<<s.grid, s.block, 0, st>>>(a); }
	// LaunchShape::buf means buffer instructions on whichever grid the shape is: the
	// cell-linear one (lin), two cells per lane (pairs) or the row grid.
	static void stream_collide(const LbmArgs& a, const LaunchShape& s, hipStream_t st) {
		LbmArgs b = a;
		b.xchunks = s.xchunks;
		b.reverse = s.alternate ? a.odd : 0u;
		if(s.lin) {
			if constexpr(linear_buffer) {
				if(s.buf) { k_stream_collide<Q, ST, TRT, EXT, linear_xcd, true, false, true><<<s.grid, s.block, s.lds, st>>>(b); return; }
			}
			k_stream_collide<Q, ST, TRT, EXT, linear_xcd, true><<<s.grid, s.block, s.lds, st>>>(b);
			return;
		}
		if constexpr(pairs) {
			if(s.pairs) {
				if constexpr(pair_buffer) {
					if(s.buf) { k_stream_collide_pair<Q, ST, TRT, EXT, true><<<s.grid, s.block, s.lds, st>>>(b); return; }
				}
				k_stream_collide_pair<Q, ST, TRT, EXT><<<s.grid, s.block, s.lds, st>>>(b);
				return;
			}
		}
		if constexpr(buffer) {
			if(s.buf) { k_stream_collide<Q, ST, TRT, EXT, xcd_map, false, false, true><<<s.grid, s.block, s.lds, st>>>(b); return; }
		}
		if constexpr(offsets32) {
			if(s.off32) { k_stream_collide<Q, ST, TRT, EXT, xcd_map, false, true><<<s.grid, s.block, s.lds, st>>>(b); return; }
		}
		k_stream_collide<Q, ST, TRT, EXT, xcd_map, false><<<s.grid, s.block, s.lds, st>>>(b);
	}
	static void update_fields(const LbmArgs& a, const LaunchShape& s, hipStream_t st) { k_update_fields<Q, ST, (EXT == 0u ? 0u : EXT_ALL)><<<s.grid, s.block, 0, st>>>(a); }
	static void update_force_field(const LbmArgs& a, const LaunchShape& s, hipStream_t st) { k_update_force_field<Q, ST><<<s.grid, s.block, 0, st>>>(a); }
	static void mass_partial(const LbmArgs& a, const LaunchShape& s, hipStream_t st, double* p) { k_mass_partial<Q, ST><<<s.grid, s.block, 0, st>>>(a, p); }
	static void update_moving_boundaries(const LbmArgs& a, const LaunchShape& s, hipStream_t st) { k_update_moving_boundaries<Q><<<s.grid, s.block, 0, st>>>(a); }
	// the adjoint kernels run on the XCD-mapped grid when rows take several blocks
	// (adjoint_shape); measured at 384x192x192: record pass +14 %, adjoint step +10 %
	static void stream_collide_record(const LbmArgs& a, const LaunchShape& s, hipStream_t st, float* m) {
		LbmArgs b = a;
		b.xchunks = s.xchunks;
		b.reverse = s.alternate ? a.odd : 0u;
		if(s.lin) k_stream_collide_record<Q, ST, TRT, EXT, false, true><<<s.grid, s.block, 0, st>>>(b, m);
		else k_stream_collide_record<Q, ST, TRT, EXT, true, false><<<s.grid, s.block, 0, st>>>(b, m);
	}
	static void adjoint_step(const AdjointArgs& g, const LaunchShape& s, hipStream_t st) {
		AdjointArgs h = g;
		h.a.xchunks = s.xchunks;
		h.a.reverse = s.alternate ? g.a.odd : 0u;
		if(s.lin) k_adjoint_step<Q, TRT, false, true><<<s.grid, s.block, 0, st>>>(h);
		else k_adjoint_step<Q, TRT, true, false><<<s.grid, s.block, 0, st>>>(h);
	}
	static void objective_partial(const LbmArgs& a, uint32_t plane_x, hipStream_t st, double* p) {
		const uint64_t cells = (uint64_t)a.Ny * a.Nz;
		k_objective_partial<Q, ST><<<(unsigned)((cells + 255u) / 256u), 256, 0, st>>>(a, plane_x, p);
	}
	// Launch tuning of stream_collide: residency cap in waves per SIMD (0 = none) and grid.
	static constexpr uint32_t waves_cap = tune_waves_cap(Q, ST, EXT);
	static constexpr bool xcd_map = tune_xcd_map(Q, ST, EXT);
	static constexpr bool linear_xcd = tune_linear_xcd(Q, ST, EXT);
	static constexpr bool pairs = tune_pairs(Q, ST, TRT, EXT);
	static constexpr bool offsets32 = tune_offsets32(Q, ST, TRT, EXT);
	static constexpr bool buffer = tune_buffer(Q, ST, TRT, EXT);
	static constexpr bool pair_buffer = tune_pair_buffer(Q, ST, TRT, EXT);
	static constexpr bool linear_buffer = tune_linear_buffer(Q, ST, TRT, EXT);
	template<int AXIS> static void transfer_axis(void* fi, void* bp, void* bm, uint32_t Nx, uint32_t Ny, uint32_t Nz, bool insert, uint32_t odd, hipStream_t st) {
		typedef typename Storage<ST>::E E;
		const uint64_t A = AXIS == 0 ? (uint64_t)Ny * Nz : AXIS == 1 ? (uint64_t)Nx * Nz : (uint64_t)Nx * Ny;
		const unsigned blocks = (unsigned)((A + 255u) / 256u);
		if(insert) k_transfer_fi<Q, E, AXIS, true><<<blocks, 256, 0, st>>>((E*)fi, (E*)bp, (E*)bm, Nx, Ny, Nz, odd);
		else k_transfer_fi<Q, E, AXIS, false><<<blocks, 256, 0, st>>>((E*)fi, (E*)bp, (E*)bm, Nx, Ny, Nz, odd);
	}
	static void transfer_fi(void* fi, void* bp, void* bm, uint32_t Nx, uint32_t Ny, uint32_t Nz, uint32_t axis, bool insert, uint32_t odd, hipStream_t st) {
		if(axis == 0u) transfer_axis<0>(fi, bp, bm, Nx, Ny, Nz, insert, odd, st);
		else if(axis == 1u) transfer_axis<1>(fi, bp, bm, Nx, Ny, Nz, insert, odd, st);
		else transfer_axis<2>(fi, bp, bm, Nx, Ny, Nz, insert, odd, st);
	}
	static const LbmKernels* table() {
		static const LbmKernels k = {initialize, stream_collide, update_fields, update_force_field, mass_partial, waves_cap,
		                             stream_collide_record, adjoint_step, objective_partial, xcd_map, transfer_fi,
		                             update_moving_boundaries, tune_linear_large(Q, ST, EXT), linear_xcd, pairs, offsets32, buffer, pair_buffer,
	                             linear_buffer};
		return &k;
	}
};

// Extension subsets with a stream_collide of their own, tried in this order; the first
// that covers the requested extensions is used (EXT_ALL covers every combination).
// Compiling the unused paths out saves registers: measured on MI355X at 256^3, the
// specialised kernel streams 3-13 % faster than the generic one.
template<int Q, uint32_t ST, uint32_t EXT> const LbmKernels* pick(bool trt) {
	return trt ? KernelSet<Q, ST, true, EXT>::table() : KernelSet<Q, ST, false, EXT>::table();
}
template<int Q, uint32_t ST> const LbmKernels* kernels_for(bool trt, uint32_t ext) {
	constexpr uint32_t E = EXT_EQUILIBRIUM_BOUNDARIES, VF = EXT_VOLUME_FORCE, FF = EXT_FORCE_FIELD;
	constexpr uint32_t POROUS = EXT_VOLUME_FORCE | EXT_POROSITY | EXT_PRECISE_DENSITY;
	if(ext == 0u) return pick<Q, ST, 0u>(trt);
	if((ext & ~E) == 0u) return pick<Q, ST, E>(trt);
	if((ext & ~VF) == 0u) return pick<Q, ST, VF>(trt);
	if((ext & ~(E | FF)) == 0u) return pick<Q, ST, E | FF>(trt);
	if((ext & ~POROUS) == 0u) return pick<Q, ST, POROUS>(trt);
	return pick<Q, ST, EXT_ALL>(trt);
}
// One translation unit per (velocity set, storage) instantiates each table.
extern template const LbmKernels* kernels_for<19, ST_FP32>(bool, uint32_t);
extern template const LbmKernels* kernels_for<19, ST_FP16S>(bool, uint32_t);
extern template const LbmKernels* kernels_for<19, ST_FP16C>(bool, uint32_t);
extern template const LbmKernels* kernels_for<27, ST_FP32>(bool, uint32_t);
extern template const LbmKernels* kernels_for<27, ST_FP16S>(bool, uint32_t);
extern template const LbmKernels* kernels_for<27, ST_FP16C>(bool, uint32_t);

inline const LbmKernels* select_kernels(uint32_t q, uint32_t storage, bool trt, uint32_t ext) {
	switch(q * 4u + storage) {
		case 19u * 4u + ST_FP32: return kernels_for<19, ST_FP32>(trt, ext);
		case 19u * 4u + ST_FP16S: return kernels_for<19, ST_FP16S>(trt, ext);
		case 19u * 4u + ST_FP16C: return kernels_for<19, ST_FP16C>(trt, ext);
		case 27u * 4u + ST_FP32: return kernels_for<27, ST_FP32>(trt, ext);
		case 27u * 4u + ST_FP16S: return kernels_for<27, ST_FP16S>(trt, ext);
		case 27u * 4u + ST_FP16C: return kernels_for<27, ST_FP16C>(trt, ext);
		default: return nullptr;
	}
}

} // namespace fx
