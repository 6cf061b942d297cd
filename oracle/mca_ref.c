/*
 * Test infrastructure, not part of the product: the reference's Minecraft
 * import loop (ImportRegionFile + WriteVoxel, Core/NBT/Importer.cpp:68-148)
 * restated over the reference's own region / NBT reader, enkiMI, which the
 * Makefile's `ref` target compiles from the reference's
 * Dependencies/enkiMI/{enkimi.c,miniz.c} (release build, NDEBUG) into
 * _ref/libmca_ref.so.  Only the import tests load it, as the checker of
 * vpt_mc_import.  The world size is a parameter (the reference fixes it at
 * 384 x 128 x 384, centred in x and z).
 */
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "enkimi.h"

#define MCA_API __attribute__((visibility("default")))

static void write_voxel(uint8_t* grid, int nx, int ny, int nz, const int32_t* origin, uint8_t v, int x, int y, int z)
{
    x = x - origin[0] + nx / 2;
    y = y - origin[1];
    z = z - origin[2] + nz / 2;
    if (y >= ny || x >= nx || z >= nz || y < 0 || x < 0 || z < 0 || v == 0) return;
    grid[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z)] = v;
}

/* One region file into grid (not cleared here).  Returns 0, or -1 when the
 * file cannot be opened; *n_sections counts the sections imported. */
MCA_API int mca_ref_import_file(const char* path, int nx, int ny, int nz, const int32_t origin[3],
                                const uint8_t lut[256], uint8_t* grid, int32_t* n_sections)
{
    FILE* fp = fopen(path, "rb");
    if (!fp) return -1;
    enkiRegionFile rf = enkiRegionFileLoad(fp);
    for (int i = 0; i < ENKI_MI_REGION_CHUNKS_NUMBER; i++) {
        enkiNBTDataStream stream;
        enkiInitNBTDataStreamForChunk(rf, i, &stream);
        if (stream.dataLength) {
            enkiChunkBlockData chunk = enkiNBTReadChunk(&stream);
            for (int s = 0; s < ENKI_MI_NUM_SECTIONS_PER_CHUNK; s++) {
                if (!chunk.sections[s]) continue;
                if (n_sections) (*n_sections)++;
                enkiMICoordinate o = enkiGetChunkSectionOrigin(&chunk, s);
                enkiMICoordinate p;
                for (p.y = 0; p.y < ENKI_MI_SIZE_SECTIONS; p.y++)
                    for (p.z = 0; p.z < ENKI_MI_SIZE_SECTIONS; p.z++)
                        for (p.x = 0; p.x < ENKI_MI_SIZE_SECTIONS; p.x++) {
                            enkiMIVoxelData v = enkiGetChunkSectionVoxelData(&chunk, s, p);
                            if (v.dataValue == 0)
                                write_voxel(grid, nx, ny, nz, origin, lut[v.blockID], o.x + p.x, o.y + p.y, o.z + p.z);
                        }
            }
        }
        enkiNBTFreeAllocations(&stream);
    }
    enkiRegionFileFreeAllocations(&rf);
    fclose(fp);
    return 0;
}

/* The reader's namespaced-id table: name and legacy id | data << 8. */
MCA_API int mca_ref_table_size(void) { return (int)enkiGetNamespaceAndBlockIDTable().size; }

MCA_API const char* mca_ref_table_name(int i)
{
    return enkiGetNamespaceAndBlockIDTable().namespaceAndBlockIDs[i].pNamespaceID;
}

MCA_API int mca_ref_table_value(int i)
{
    const enkiMINamespaceAndBlockID* e = &enkiGetNamespaceAndBlockIDTable().namespaceAndBlockIDs[i];
    return (int)e->blockID | ((int)e->dataValue << 8);
}
