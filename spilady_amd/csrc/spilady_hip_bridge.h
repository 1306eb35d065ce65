// spilady_hip_bridge.h - declarations a SPILADY build needs to run on
// libspilady_hip.so instead of its CUDA files.
//
// SPILADY declares its GPU lifecycle functions only under -DGPU
// (prototype_CPU.h:275-289), and -DGPU also pulls in the CUDA random-number
// headers (spilady.h:38-41). A build on this library defines HIPLIB instead
// and force-includes this header (g++ -include spilady_hip_bridge.h), which
// declares the same functions with the same signatures. Their definitions are
// in spilady_hip_bridge.cpp.
#ifndef SPILADY_HIP_BRIDGE_H
#define SPILADY_HIP_BRIDGE_H

#if defined HIPLIB && !defined GPU
void initial_GPU();
void potential_table_GPU();
void free_potential_ptr_GPU();
void copy_CPU_to_GPU();
void copy_atoms_from_GPU_to_CPU();
void copy_cells_from_GPU_to_CPU();
void free_copy_CPU_to_GPU();
#endif

#endif
