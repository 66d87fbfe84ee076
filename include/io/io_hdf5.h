/*
 * Field output, same entry points as the reference's src/io/io_hdf5.h:1-52
 * (set_name_*, io_hdf5_dump_mesh, io_hdf5_dump_data, io_hdf5_write_xdmf,
 * io_save_all), so that its drivers build unchanged. HDF5 is not required:
 * the same datasets are written as raw little-endian binary files and the
 * XDMF file (Tet_10 topology, node geometry, one attribute per macro field)
 * points into them with Format="Binary" data items (ParaView, VisIt).
 *
 *  - mesh file  <basename>_mesh.bin : nb_elems (int32), then
 *    connectivity_000000 (int32, nb_elems x 10, elem2qnode), then
 *    nodes_000000 (float64, nbqnodes_in x 3);
 *  - data file  <prefix>.bin : f00000 ... f<m-1> (float64, nbqnodes_in each),
 *    the value of component m at each global node taken from the first
 *    element of its node2elem list (as src/io/io_hdf5.c:148-205);
 *  - <basename>_simulation_<iternb>.xmf ; kinetic schemes get w<iw> as the
 *    sum over velocities (XDMF Function items), like the reference.
 */
#ifndef IO_HDF5_H
#define IO_HDF5_H

#include <simulation/simulation.h>

#ifdef __cplusplus
extern "C" {
#endif

void set_name_mesh(char *filename, char *basename);
void set_name_data(char *filename, char *basename, int iternb);
void set_name_xmf(char *filename, char *basename, int iternb);
void io_hdf5_dump_mesh(gdn_simulation *tf, char *geo_filename);
void io_hdf5_dump_data(gdn_simulation *tf, char *data_filename_prefix);
void io_hdf5_write_xdmf(gdn_simulation *tf, char *filename, char *geo_filename,
			char *data_filename_prefix);
void io_save_all(gdn_simulation *simulation, char filename[1024], int iternb);

#ifdef __cplusplus
}
#endif
#endif
