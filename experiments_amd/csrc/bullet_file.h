// Reader of Bullet's serialized scene files (".bullet", the format the
// reference loads through bParse::btBulletFile in
// opencl/gpu_rigidbody_pipeline2/BulletDataExtractor.cpp:454-520).
//
// A file is a 12-byte header, a sequence of chunks and a DNA chunk that
// describes every struct layout in the file.  Nothing here assumes a layout:
// field offsets come from the file's own DNA, so files written with 4- or
// 8-byte pointers and either byte order read the same way.  Host code only.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

namespace bullet {

// Bullet's shape type numbers (btBroadphaseProxy.h BroadphaseNativeTypes),
// the ones the loader distinguishes.
enum ShapeType {
    BOX = 0,
    CONVEX_HULL = 4,
    SPHERE = 8,
    TRIANGLE_MESH = 21,
    STATIC_PLANE = 28,
    COMPOUND = 31,
};

struct Field {
    std::string name;   // without '*' and array suffix
    int type;           // index into File::types
    bool pointer;       // '*name' or '(*name)()'
    int count;          // array elements (1 if not an array)
    size_t offset;      // from the start of the struct
    size_t size;        // bytes of the whole member
};

struct Struct {
    int type;           // index into File::types
    size_t size;        // bytes (DNA TLEN, checked against the fields)
    std::vector<Field> fields;
};

struct Chunk {
    uint32_t code;      // four characters, as stored
    int32_t len;        // bytes of data
    uint64_t old_ptr;   // address the data had when written
    int32_t sdna;       // struct index of the elements
    int32_t nr;         // element count
    size_t data;        // file offset of the data
};

struct File {
    std::vector<uint8_t> bytes;
    bool doubles = false;     // 'BULLETd': double-precision scalars
    int ptr_size = 4;
    bool swap = false;        // big-endian file
    int version = 0;
    std::vector<std::string> names, types;
    std::vector<int> tlen;
    std::vector<Struct> structs;
    std::unordered_map<std::string, int> struct_of_type;  // type name -> struct index
    std::vector<Chunk> chunks;
    std::unordered_map<uint64_t, size_t> chunk_of_ptr;    // old pointer -> chunk index

    // Parse; false with a message on a malformed or truncated file.
    bool parse(std::vector<uint8_t> data, std::string &err);

    // A value inside element `elem` of a chunk, by a field path such as
    // "m_collisionObjectData.m_worldTransform.m_origin.m_floats" and an
    // index into the last field's array.  False if the path does not exist
    // in this file's DNA or lies outside the chunk.
    bool read_float(const Chunk &c, int elem, const char *path, int index, float &out) const;
    bool read_int(const Chunk &c, int elem, const char *path, int index, int32_t &out) const;
    bool read_ptr(const Chunk &c, int elem, const char *path, uint64_t &out) const;
    // chunk a pointer refers to, or nullptr
    const Chunk *deref(uint64_t ptr) const;
    const char *type_of(const Chunk &c) const;

  private:
    bool locate(const Chunk &c, int elem, const char *path, int index, const Field **f,
                size_t &at) const;
    uint32_t u32(size_t at) const;
    uint16_t u16(size_t at) const;
    uint64_t u64(size_t at) const;
};

// What the reference's extractor takes from a file: collision shapes in
// file order and rigid bodies with the index of their shape.
struct ShapeRecord {
    int type = -1;
    float local_scaling[3] = {1, 1, 1};
    float implicit_dims[3] = {0, 0, 0};  // box: half extents; sphere: radius in [0]
    float margin = 0;
    float plane_normal[3] = {0, 0, 0};
    float plane_constant = 0;
    std::vector<float> points;           // convex hull: unscaled xyz
    uint64_t old_ptr = 0;
};
struct BodyRecord {
    int shape = -1;                      // index into the shape list, -1 if none
    float inverse_mass = 0;
    float basis[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};  // rows
    float origin[3] = {0, 0, 0};
    float linvel[3] = {0, 0, 0};
    float angvel[3] = {0, 0, 0};
    float friction = 0, restitution = 0;
};
struct Scene {
    std::vector<ShapeRecord> shapes;
    std::vector<BodyRecord> bodies;
};
bool extract(const File &f, Scene &out, std::string &err);

// Rotation matrix (rows) to a unit quaternion (x, y, z, w), w >= 0 when the
// trace is positive (Shepperd's method: the largest of w, x, y, z is formed
// from the diagonal, the others from off-diagonal sums and differences).
void basis_to_quat(const float m[9], float q[4]);

}  // namespace bullet
