// C++ façade over the C ABI, mirroring the reference's VoxelRT::World
// (Core/World.h:35-222): the block grid stays on the host as m_WorldData,
// the device copy and its distance field live in HBM.  Header-only; links
// against libvpt.so.
#pragma once

#include <stdint.h>

#include <array>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "vpt.h"

namespace VoxelRT {

struct Block {
    uint8_t block;
};

inline void vpt_check(int rc)
{
    if (rc != VPT_OK) throw std::runtime_error(std::string("vpt: ") + vpt_last_error());
}

class Context {
public:
    explicit Context(int device = 0) { vpt_check(vpt_ctx_create(device, &m_Ctx)); }
    ~Context() { vpt_ctx_destroy(m_Ctx); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    vpt_ctx* Handle() const { return m_Ctx; }

private:
    vpt_ctx* m_Ctx = nullptr;
};

class World {
public:
    World(Context& ctx, int nx = 384, int ny = 128, int nz = 384)
        : m_X(nx), m_Y(ny), m_Z(nz), m_WorldData((size_t)nx * ny * nz, Block{0})
    {
        vpt_check(vpt_world_create(ctx.Handle(), nx, ny, nz, &m_World));
    }
    ~World() { vpt_world_destroy(m_World); }
    World(const World&) = delete;
    World& operator=(const World&) = delete;

    // World.h:46-70
    const Block& GetBlock(uint16_t x, uint16_t y, uint16_t z) const { return m_WorldData[Index(x, y, z)]; }
    void SetBlock(uint16_t x, uint16_t y, uint16_t z, Block block)
    {
        m_WorldData[Index(x, y, z)] = block;
        m_Pending[(uint32_t)x | ((uint32_t)y << 10) | ((uint32_t)z << 20)] = block.block;
    }

    // World.h:167-171: upload the whole grid
    void Buffer(void* stream = nullptr)
    {
        vpt_check(vpt_world_upload(m_World, reinterpret_cast<const uint8_t*>(m_WorldData.data()), stream));
        m_Pending.clear();
    }

    // MCWorldImporter::ImportWorld (NBT/Importer.cpp:150-168): the region
    // files of `dir` into m_WorldData, then Buffer(); rebuild the DF after it
    // (Pipeline.cpp:1236, 1293-1294).  names: palette name -> legacy id.
    void ImportMinecraftWorld(const char* dir, const float origin[3], const uint8_t mc_to_block[256],
                              const vpt_mc_block* names = nullptr, int n_names = 0, void* stream = nullptr)
    {
        vpt_check(vpt_mc_import(dir, m_X, m_Y, m_Z, origin, mc_to_block, names, n_names,
                                reinterpret_cast<uint8_t*>(m_WorldData.data()), nullptr));
        Buffer(stream);
    }

    // World.h:173-174
    void InitializeDistanceGenerator() {}
    void GenerateDistanceField(void* stream = nullptr)
    {
        FlushEdits(stream);
        vpt_check(vpt_df_build(m_World, stream));
    }

    // Pending edits flushed, then the DF updated around them only (the
    // reference follows SetBlock with a full GenerateDistanceField,
    // World.cpp:367-375); asynchronous on `stream`.
    void UpdateDistanceField(void* stream = nullptr)
    {
        if (m_Pending.empty()) return;
        std::vector<uint32_t> e;
        e.reserve(m_Pending.size());
        for (auto& kv : m_Pending) e.push_back(kv.first);
        FlushEdits(stream);
        vpt_check(vpt_df_update(m_World, e.data(), (int)e.size(), nullptr, stream));
    }

    // World.h:186: CPU block picking
    std::array<int, 4> RaycastDetect(const float pos[3], const float dir[3]) const
    {
        std::array<int32_t, 4> out{};
        vpt_check(vpt_raycast_detect(reinterpret_cast<const uint8_t*>(m_WorldData.data()), m_X, m_Y, m_Z, pos, dir,
                                     out.data()));
        return {out[0], out[1], out[2], out[3]};
    }

    // Pending SetBlock edits as one batched upload (the reference issues a
    // one-voxel glTexSubImage3D per edit, World.cpp:367-374).
    void FlushEdits(void* stream = nullptr)
    {
        if (m_Pending.empty()) return;
        std::vector<uint32_t> e;
        std::vector<uint8_t> ids;
        for (auto& kv : m_Pending) {
            e.push_back(kv.first);
            ids.push_back(kv.second);
        }
        vpt_check(vpt_world_set_blocks(m_World, e.data(), ids.data(), (int)e.size(), stream));
        m_Pending.clear();
    }

    vpt_world* Handle() const { return m_World; }
    int SizeX() const { return m_X; }
    int SizeY() const { return m_Y; }
    int SizeZ() const { return m_Z; }

    std::vector<Block> m_WorldData;

private:
    size_t Index(int x, int y, int z) const { return (size_t)x + (size_t)y * m_X + (size_t)z * m_X * m_Y; }
    int m_X, m_Y, m_Z;
    vpt_world* m_World = nullptr;
    std::unordered_map<uint32_t, uint8_t> m_Pending;
};

}  // namespace VoxelRT
