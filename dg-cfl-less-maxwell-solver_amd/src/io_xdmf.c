/*
 * Field output as raw binary + XDMF (include/io/io_hdf5.h): the datasets of
 * the reference's HDF5 writer (src/io/io_hdf5.c:60-368) in plain files.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <io/io_hdf5.h>
#include <mesh/t10.h>

#include "host_internal.h"

void set_name_mesh(char *filename, char *basename)
{
	snprintf(filename, 1024, "%s_mesh.bin", basename);
}

void set_name_data(char *filename, char *basename, int iternb)
{
	snprintf(filename, 1024, "%s_data_%04d", basename, iternb);
}

void set_name_xmf(char *filename, char *basename, int iternb)
{
	snprintf(filename, 1024, "%s_simulation_%04d.xmf", basename, iternb);
}

static FILE *open_or_die(const char *name, const char *mode)
{
	FILE *f = fopen(name, mode);
	if (!f)
		gdn_fatal("cannot open an output file");
	return f;
}

void io_hdf5_dump_mesh(gdn_simulation *simu, char *geo_filename)
{
	const gdn_mesh *m = simu->tt;
	FILE *f = open_or_die(geo_filename, "wb");
	const int32_t ne = m->nbelems;
	fwrite(&ne, sizeof(ne), 1, f);
	int32_t *conn = malloc(sizeof(int32_t) * 10 * (size_t)ne);
	if (!conn)
		gdn_fatal("out of memory");
	for (long k = 0; k < 10L * ne; k++)
		conn[k] = m->elem2qnode[k];
	fwrite(conn, sizeof(int32_t), 10 * (size_t)ne, f);
	free(conn);
	fwrite(m->node, sizeof(double), 3 * (size_t)m->nbqnodes_in, f);
	fclose(f);
}

void io_hdf5_dump_data(gdn_simulation *simu, char *data_filename_prefix)
{
	const gdn_mesh *t = simu->tt;
	const int np = t->nbqnodes_in, m = simu->m;
	char name[1100];
	snprintf(name, sizeof(name), "%s.bin", data_filename_prefix);
	double *value = malloc(sizeof(double) * (size_t)m * np);
	if (!value)
		gdn_fatal("out of memory");
	for (int in = 0; in < np; in++) {
		const int ie = t->node2elem[t->max_node2elem * in];
		int iloc = 0;
		while (t->elem2qnode[ie * NQN + iloc] != in)
			iloc++;
		const long ipg = (long)ie * NQN + iloc;
		for (int im = 0; im < m; im++)
			value[(size_t)im * np + in] = simu->use_kinetic_scheme ?
							      simu->wn[(long)im * simu->neq + ipg] :
							      simu->wn[ipg * m + im];
	}
	FILE *f = open_or_die(name, "wb");
	fwrite(value, sizeof(double), (size_t)m * np, f);
	fclose(f);
	free(value);
}

/* file name without directories: XDMF resolves it next to the .xmf file */
static const char *base_name(const char *p)
{
	const char *s = strrchr(p, '/');
	return s ? s + 1 : p;
}

static void data_item(FILE *x, const char *file, long seek, int n, int cols,
		      const char *type, int prec)
{
	fprintf(x, "<DataItem Format=\"Binary\" Endian=\"Little\" NumberType=\"%s\" "
		   "Precision=\"%d\" Seek=\"%ld\" Dimensions=\"%d %d\">%s</DataItem>\n",
		type, prec, seek, n, cols, base_name(file));
}

void io_hdf5_write_xdmf(gdn_simulation *simu, char *filename, char *geo_filename,
			char *data_filename_prefix)
{
	const gdn_mesh *t = simu->tt;
	const int np = t->nbqnodes_in, ne = t->nbelems;
	char data[1100];
	snprintf(data, sizeof(data), "%s.bin", data_filename_prefix);
	FILE *x = open_or_die(filename, "w");
	fprintf(x, "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n");
	fprintf(x, "<Xdmf Version=\"2.0\">\n<Domain>\n<Grid Name=\"TheMesh\">\n");
	fprintf(x, "<Time Value=\"%.17g\" />\n", simu->t);
	fprintf(x, "<Topology TopologyType=\"Tet_10\" NumberOfElements=\"%d\">\n", ne);
	data_item(x, geo_filename, 4, ne, 10, "Int", 4);
	fprintf(x, "</Topology>\n<Geometry GeometryType=\"XYZ\">\n");
	data_item(x, geo_filename, 4 + 40L * ne, np, 3, "Float", 8);
	fprintf(x, "</Geometry>\n");
	if (simu->use_kinetic_scheme) {
		const int nv = simu->mdl->nb_v;
		for (int iw = 0; iw < simu->mdl->nb_w; iw++) {
			fprintf(x, "<Attribute Name=\"w%d\" Center=\"Node\">\n", iw);
			fprintf(x, "<DataItem ItemType=\"Function\" Function=\"");
			for (int iv = 0; iv < nv; iv++)
				fprintf(x, "%s$%d", iv ? "+" : "", iv);
			fprintf(x, "\" Dimensions=\"%d 1\">\n", np);
			for (int iv = 0; iv < nv; iv++)
				data_item(x, data, 8L * np * (iw * nv + iv), np, 1, "Float", 8);
			fprintf(x, "</DataItem>\n</Attribute>\n");
		}
	} else {
		for (int iv = 0; iv < simu->m; iv++) {
			fprintf(x, "<Attribute Name=\"w%d\" Center=\"Node\">\n", iv);
			data_item(x, data, 8L * np * iv, np, 1, "Float", 8);
			fprintf(x, "</Attribute>\n");
		}
	}
	fprintf(x, "</Grid>\n</Domain>\n</Xdmf>\n");
	fclose(x);
}

void io_save_all(gdn_simulation *simulation, char filename[1024], int iternb)
{
	char mesh_filename[1024] = { 0 };
	char data_filename[1024] = { 0 };
	char xmf_filename[1024] = { 0 };
	set_name_mesh(mesh_filename, filename);
	set_name_data(data_filename, filename, iternb);
	set_name_xmf(xmf_filename, filename, iternb);
	io_hdf5_dump_mesh(simulation, mesh_filename);
	io_hdf5_dump_data(simulation, data_filename);
	io_hdf5_write_xdmf(simulation, xmf_filename, mesh_filename, data_filename);
}
