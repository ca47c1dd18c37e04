#pragma once
// Compile-time switches of a scene, with the reference's names (src/defines.hpp:5-25).
// Edit this file like the reference's defines.hpp, or pass -D<SWITCH> to the compiler.
// All kernel variants are precompiled for gfx950; the switches only select which one the
// LBM class asks the C ABI for.

#if !defined(D3Q19) && !defined(D3Q27)
#define D3Q19 // D3Q19 velocity set; 93 (FP32) or 55 (FP16) bytes/cell
//#define D3Q27 // D3Q27 velocity set; 125 (FP32) or 71 (FP16) bytes/cell
#endif

#if !defined(SRT) && !defined(TRT)
#define SRT // single-relaxation-time collision
//#define TRT // two-relaxation-time collision (magic parameter 3/16)
#endif

//#define FP16S // DDFs stored as range-shifted IEEE binary16
//#define FP16C // DDFs stored as the custom 1-4-11 format

//#define BENCHMARK // run the benchmark scene with all extensions disabled

//#define VOLUME_FORCE // global force per volume (fx,fy,fz) with Guo forcing
//#define FORCE_FIELD // per-cell force field F and lbm.update_force_field()/object_force()
//#define EQUILIBRIUM_BOUNDARIES // TYPE_E cells with preset density/velocity
//#define UPDATE_FIELDS // rho/u written by every stream_collide
//#define POROSITY // per-cell Brinkman porosity lbm.porosity (not in the reference)
//#define PRECISE_DENSITY // equilibrium takes rho-1 as the DDF sum (not in the reference; FP32 accuracy at tiny forces)
//#define MOVING_BOUNDARIES // TYPE_S cells with a velocity move (TYPE_MS marks their neighbours)
//#define SUBGRID // Smagorinsky-Lilly subgrid turbulence model

#define TYPE_S 0x01 // (stationary) solid boundary
#define TYPE_E 0x02 // equilibrium boundary (inflow/outflow)
#define TYPE_MS 0x03 // cell next to a moving solid boundary (MOVING_BOUNDARIES)
#define TYPE_T 0x04 // temperature boundary (no TEMPERATURE extension in this build)
#define TYPE_F 0x08 // fluid (SURFACE, not in this build)
#define TYPE_I 0x10 // interface (SURFACE, not in this build)
#define TYPE_G 0x20 // gas (SURFACE, not in this build)
#define TYPE_X 0x40 // reserved type X
#define TYPE_Y 0x80 // reserved type Y

#ifdef BENCHMARK
#undef UPDATE_FIELDS
#undef VOLUME_FORCE
#undef FORCE_FIELD
#undef EQUILIBRIUM_BOUNDARIES
#undef POROSITY
#undef PRECISE_DENSITY
#undef MOVING_BOUNDARIES
#undef SUBGRID
#endif

#if defined(FP16S) && defined(FP16C)
#error "choose at most one of FP16S and FP16C"
#endif
#if defined(D3Q19) && defined(D3Q27)
#error "choose one velocity set"
#endif
#if defined(SRT) && defined(TRT)
#error "choose one collision operator"
#endif
#if defined(SURFACE) || defined(TEMPERATURE) || defined(PARTICLES) || defined(GRAPHICS) || defined(INTERACTIVE_GRAPHICS) || defined(INTERACTIVE_GRAPHICS_ASCII)
#error "this build implements the LBM core (VOLUME_FORCE, FORCE_FIELD, EQUILIBRIUM_BOUNDARIES, MOVING_BOUNDARIES, UPDATE_FIELDS, SUBGRID, POROSITY); the requested extension is not available"
#endif

#if defined(FP16S) || defined(FP16C)
#define fpxx unsigned short
#else
#define fpxx float
#endif
