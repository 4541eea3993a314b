/*
 * generativeqc.h -- context / system / status C ABI of the MI355X DF-JK engine.
 *
 * This header declares the subset of the reference's public C ABI that the
 * density-fitted Fock path binds (include/generativeqc/generativeqc.h in the
 * reference).  Names, struct layouts and numeric status values are kept
 * identical so an existing ctypes/cffi binding of the reference can load this
 * library unchanged for these entry points.  Each declaration cites the
 * reference declaration it replaces.
 */
#ifndef GENERATIVEQC_GENERATIVEQC_H
#define GENERATIVEQC_GENERATIVEQC_H

#include <stddef.h>
#include <stdint.h>

#define GENERATIVEQC_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

/* Reference: generativeqc.h:21 */
#define GENERATIVEQC_ABI_VERSION 0u

/* Reference: generativeqc.h:23-42 (same numeric values). */
typedef int32_t generativeqc_status;
enum {
  GENERATIVEQC_STATUS_SUCCESS = 0,
  GENERATIVEQC_STATUS_INVALID_ARGUMENT = 1,
  GENERATIVEQC_STATUS_ABI_MISMATCH = 2,
  GENERATIVEQC_STATUS_NOT_IMPLEMENTED = 3,
  GENERATIVEQC_STATUS_NOT_CONVERGED = 4,
  GENERATIVEQC_STATUS_SCF_NOT_CONVERGED = GENERATIVEQC_STATUS_NOT_CONVERGED,
  GENERATIVEQC_STATUS_NUMERICAL_FAILURE = 5,
  /* Device runtime failure.  The reference names it CUDA_ERROR; here it is
   * any HIP runtime, rocBLAS or rocSOLVER failure on the MI355X device. */
  GENERATIVEQC_STATUS_CUDA_ERROR = 6,
  GENERATIVEQC_STATUS_DEVICE_ERROR = GENERATIVEQC_STATUS_CUDA_ERROR,
  GENERATIVEQC_STATUS_OUT_OF_MEMORY = 7,
  GENERATIVEQC_STATUS_INTERNAL_ERROR = 8,
  GENERATIVEQC_STATUS_PRECISION_UNAVAILABLE = 9
};

/* Reference: generativeqc.h:60-65.  Value 1 selects the accelerator; in this
 * library that is the HIP device (gfx950).  CPU_REFERENCE is not provided:
 * creating such a context returns NOT_IMPLEMENTED (there is no CPU path). */
typedef int32_t generativeqc_backend;
enum {
  GENERATIVEQC_BACKEND_CPU_REFERENCE = 0,
  GENERATIVEQC_BACKEND_CUDA = 1,
  GENERATIVEQC_BACKEND_HIP = GENERATIVEQC_BACKEND_CUDA,
  GENERATIVEQC_BACKEND_HYBRID_CUDA = 2
};

/* Reference: generativeqc.h:108-113. */
typedef int32_t generativeqc_basis_representation;
enum { GENERATIVEQC_BASIS_CARTESIAN = 0, GENERATIVEQC_BASIS_SPHERICAL = 1 };

typedef struct generativeqc_context generativeqc_context;
typedef struct generativeqc_system generativeqc_system;

/* Reference: generativeqc.h:560-565. */
typedef struct generativeqc_context_descriptor {
  uint32_t struct_size;
  uint32_t abi_version;
  int32_t device_id;
  generativeqc_backend backend;
} generativeqc_context_descriptor;

/* Reference: generativeqc.h:567-600 (coordinates in Bohr). */
typedef struct generativeqc_atom {
  int32_t atomic_number;
  double x;
  double y;
  double z;
} generativeqc_atom;

typedef struct generativeqc_primitive {
  double exponent;
  double coefficient;
} generativeqc_primitive;

typedef struct generativeqc_shell {
  uint32_t atom_index;
  uint32_t angular_momentum;
  uint32_t primitive_offset;
  uint32_t primitive_count;
} generativeqc_shell;

typedef struct generativeqc_system_descriptor {
  uint32_t struct_size;
  uint32_t abi_version;
  const generativeqc_atom* atoms;
  uint32_t atom_count;
  const generativeqc_shell* shells;
  uint32_t shell_count;
  const generativeqc_primitive* primitives;
  uint32_t primitive_count;
  int32_t charge;
  uint32_t multiplicity;
  generativeqc_basis_representation basis_representation;
} generativeqc_system_descriptor;

/* Reference: generativeqc.h:1186. */
GENERATIVEQC_API uint32_t generativeqc_get_abi_version(void);
/* Reference: generativeqc.h:1189. */
GENERATIVEQC_API const char* generativeqc_get_source_identity(void);
/* Reference: generativeqc.h:1215. */
GENERATIVEQC_API const char* generativeqc_status_message(generativeqc_status status);

/* Reference: generativeqc.h:1225-1229.  Requires backend CUDA/HIP; binds the
 * HIP device `device_id` and owns one HIP stream plus rocBLAS/rocSOLVER state. */
GENERATIVEQC_API generativeqc_status generativeqc_context_create(
    const generativeqc_context_descriptor* descriptor, generativeqc_context** context);
GENERATIVEQC_API void generativeqc_context_destroy(generativeqc_context* context);
/* Reference: generativeqc.h:1231-1235. */
GENERATIVEQC_API const char* generativeqc_context_get_last_detail(const generativeqc_context* context);
GENERATIVEQC_API const char* generativeqc_context_last_error(const generativeqc_context* context);

/* Reference: generativeqc.h:1237-1240.  Copies and normalises the basis
 * (src/molecule/basis.cpp:186-282); the descriptor may be freed afterwards. */
GENERATIVEQC_API generativeqc_status generativeqc_system_create(
    generativeqc_context* context, const generativeqc_system_descriptor* descriptor,
    generativeqc_system** system);
GENERATIVEQC_API void generativeqc_system_destroy(generativeqc_system* system);

#ifdef __cplusplus
}
#endif
#endif
