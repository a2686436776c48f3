/* mi_rb_clphysicsdemo.hpp -- header-only C++ class with the method names,
 * argument meaning and call order of the reference's CLPhysicsDemo
 * (opencl/gpu_rigidbody_pipeline2/CLPhysicsDemo.h:31-78), forwarding to the
 * mi_rb C ABI.  A demo that includes this header instead of CLPhysicsDemo.h
 * and links libmi_rb.so runs its step on the gfx950 kernels.
 *
 * Differences a caller can see:
 *   - getBodiesGpu() returns a non-owning mi_rb_device_view pointer instead of
 *     a cl_mem (the bodies are structure-of-arrays in device memory);
 *   - a concave mesh takes a shape slot per triangle (plus one for the
 *     mesh) and collides with convex hull bodies only;
 *   - errors are reported by last_status() (mi_rb_status of the last call)
 *     instead of asserts and printf.
 */
#ifndef MI_RB_CLPHYSICSDEMO_HPP
#define MI_RB_CLPHYSICSDEMO_HPP

#include <vector>

#include "mi_rb.h"

struct CLPhysicsDemo {
    int m_numPhysicsInstances = 0;
    int m_numDynamicPhysicsInstances = 0;
    int m_maxShapeBufferCapacityInBytes = 0;
    int m_maxNumPhysicsInstances = 0;

    /* CLPhysicsDemo(int maxShapeBufferCapacity, int maxNumObjects) */
    CLPhysicsDemo(int maxShapeBufferCapacity, int maxNumObjects)
        : m_maxShapeBufferCapacityInBytes(maxShapeBufferCapacity),
          m_maxNumPhysicsInstances(maxNumObjects) {
        mi_rb_config_reference(&m_cfg);
        m_cfg.max_bodies = maxNumObjects;
    }
    virtual ~CLPhysicsDemo() { cleanup(); }

    /* init(preferredDevice, preferredPlatform, useInterop): the device index
     * selects the GPU; platform and interop have no meaning here */
    void init(int preferredDevice, int preferredPlatform, bool useInterop) {
        (void)preferredPlatform;
        (void)useInterop;
        if (preferredDevice >= 0) m_cfg.device = preferredDevice;
        m_status = mi_rb_create(&m_cfg, &m_world);
    }

    /* the bodies in device memory (replaces the cl_mem of the body array) */
    const mi_rb_device_view *getBodiesGpu() {
        m_status = mi_rb_device_ptrs(m_world, &m_view);
        return m_status == MI_RB_OK ? &m_view : nullptr;
    }

    int registerConvexPolyhedron(const float *vertices, int strideInBytes, int numVertices,
                                 const float *scaling, bool noHeightField) {
        (void)noHeightField;
        int shape = -1;
        m_status = mi_rb_register_convex_hull(m_world, vertices, strideInBytes, numVertices,
                                              scaling, &shape);
        return m_status == MI_RB_OK ? shape : -1;
    }
    /* registerConvexPolyhedron(btConvexUtility* utilPtr, bool noHeightField)
     * (CLPhysicsDemo.cpp:331-410): registers the caller's polyhedral
     * features as they are, in their order.  Util: anything with the members
     * of btConvexUtility that the reference reads -- m_vertices (points with
     * operator[] giving x, y, z), m_faces (m_plane[4], m_indices) and
     * m_uniqueEdges -- each a container with size() and operator[]. */
    template <class Util>
    int registerConvexPolyhedron(Util *utilPtr, bool noHeightField) {
        (void)noHeightField;
        const int nv = (int)utilPtr->m_vertices.size(), nf = (int)utilPtr->m_faces.size();
        const int ne = (int)utilPtr->m_uniqueEdges.size();
        std::vector<float> verts(3 * (size_t)nv), planes(4 * (size_t)nf), edges(3 * (size_t)ne);
        std::vector<int> face_index(2 * (size_t)nf), indices;
        for (int i = 0; i < nv; i++)
            for (int k = 0; k < 3; k++) verts[3 * i + k] = utilPtr->m_vertices[i][k];
        for (int f = 0; f < nf; f++) {
            const auto &face = utilPtr->m_faces[f];
            for (int k = 0; k < 4; k++) planes[4 * f + k] = face.m_plane[k];
            face_index[2 * f] = (int)indices.size();
            face_index[2 * f + 1] = (int)face.m_indices.size();
            for (int k = 0; k < (int)face.m_indices.size(); k++) indices.push_back(face.m_indices[k]);
        }
        for (int i = 0; i < ne; i++)
            for (int k = 0; k < 3; k++) edges[3 * i + k] = utilPtr->m_uniqueEdges[i][k];
        int shape = -1;
        m_status = mi_rb_register_convex_polyhedron(
            m_world, verts.data(), 3 * (int)sizeof(float), nv, planes.data(), face_index.data(), nf,
            indices.data(), (int)indices.size(), edges.data(), 3 * (int)sizeof(float), ne, &shape);
        return m_status == MI_RB_OK ? shape : -1;
    }
    int registerSphereShape(float radius) {
        int shape = -1;
        m_status = mi_rb_register_sphere(m_world, radius, &shape);
        return m_status == MI_RB_OK ? shape : -1;
    }
    int registerPlaneShape(const float planeNormal[3], float planeConstant) {
        int shape = -1;
        m_status = mi_rb_register_plane(m_world, planeNormal, planeConstant, &shape);
        return m_status == MI_RB_OK ? shape : -1;
    }
    /* childShapes: a container (size(), operator[]) of btGpuChildShape-like
     * records: float m_childPosition[4], m_childOrientation[4], int m_shapeIndex */
    template <class Children>
    int registerCompoundShape(Children *childShapes) {
        const int n = (int)childShapes->size();
        std::vector<int> shapes(n);
        std::vector<float> pos(3 * (size_t)n), orn(4 * (size_t)n);
        for (int i = 0; i < n; i++) {
            const auto &c = (*childShapes)[i];
            shapes[i] = c.m_shapeIndex;
            for (int k = 0; k < 3; k++) pos[3 * i + k] = c.m_childPosition[k];
            for (int k = 0; k < 4; k++) orn[4 * i + k] = c.m_childOrientation[k];
        }
        int shape = -1;
        m_status = mi_rb_register_compound(m_world, n, shapes.data(), pos.data(), orn.data(), &shape);
        return m_status == MI_RB_OK ? shape : -1;
    }

    /* registerConcaveMesh(vertices, indices, scaling) (CLPhysicsDemo.cpp:287-300):
     * Vertices: an array of points with operator[] giving x, y, z (e.g.
     * btAlignedObjectArray<btVector3>); Indices: ints, three per triangle */
    template <class Vertices, class Indices>
    int registerConcaveMesh(Vertices *vertices, Indices *indices, const float *scaling) {
        const int nv = (int)vertices->size(), ni = (int)indices->size();
        std::vector<float> xyz(3 * (size_t)nv);
        std::vector<int> ix(ni);
        for (int i = 0; i < nv; i++)
            for (int k = 0; k < 3; k++) xyz[3 * i + k] = (*vertices)[i][k];
        for (int i = 0; i < ni; i++) ix[i] = (*indices)[i];
        int shape = -1;
        m_status = mi_rb_register_concave_mesh(m_world, xyz.data(), 3 * (int)sizeof(float), nv,
                                               ix.data(), ni / 3, scaling, &shape);
        return m_status == MI_RB_OK ? shape : -1;
    }

    /* position: float[3] (or [4]), orientation: quaternion xyzw */
    int registerPhysicsInstance(float mass, const float *position, const float *orientation,
                                int collisionShapeIndex, int userPointer) {
        (void)userPointer;
        int body = -1;
        m_status = mi_rb_register_body(m_world, mass, position, orientation, collisionShapeIndex,
                                       0, &body);
        if (m_status != MI_RB_OK) return -1;
        m_numPhysicsInstances++;
        if (mass != 0.f) m_numDynamicPhysicsInstances++;
        return body;
    }

    void setObjectTransform(const float *position, const float *orientation, int objectIndex) {
        m_status = mi_rb_set_transform(m_world, objectIndex, position, orientation);
    }
    void setObjectLinearVelocity(const float *velocity, int objectIndex) {
        m_status = mi_rb_set_linvel(m_world, objectIndex, velocity);
    }

    /* velocities set above are written through at once */
    void writeVelocitiesToGpu() { m_status = MI_RB_OK; }
    /* shapes, bodies and every derived table go to the device */
    void writeBodiesToGpu() { m_status = mi_rb_upload(m_world); }

    void readbackBodiesToCpu() { m_status = mi_rb_readback(m_world); }
    void getObjectTransformFromCpu(float *position, float *orientation, int objectIndex) {
        m_status = mi_rb_get_transform(m_world, objectIndex, position, orientation);
    }

    void cleanup() {
        if (m_world) mi_rb_destroy(m_world);
        m_world = nullptr;
    }

    void stepSimulation() { m_status = mi_rb_step(m_world, 1); }

    /* additions: the status of the last call, and the world handle for the
     * rest of the C ABI (statistics, graph replay, warm start) */
    mi_rb_status last_status() const { return m_status; }
    void *world() const { return m_world; }

  private:
    friend void copyTransformsToBVO(CLPhysicsDemo &demo, void *vbo, int positionOffset,
                                    int vboFloat4Capacity);
    friend bool createSceneFromBulletFile(CLPhysicsDemo &demo, const char *fileName,
                                          bool keepStaticObjects);
    mi_rb_config m_cfg{};
    void *m_world = nullptr;
    mi_rb_device_view m_view{};
    mi_rb_status m_status = MI_RB_OK;
};

/* copyTransformsToBVO (opencl/broadphase_benchmark/findPairsOpenCL.cpp:140-160), the
 * demo's render interop step (gpu_rigidbody_pipeline2/main.cpp:274-278): the
 * reference takes its btFindPairsIO and the body cl_mem; here the demo and
 * the device pointer of the vertex buffer (mapped for HIP by the renderer),
 * the positions' offset in floats and the buffer's float4 capacity. */
inline void copyTransformsToBVO(CLPhysicsDemo &demo, void *vbo, int positionOffset,
                                int vboFloat4Capacity) {
    demo.m_status = mi_rb_copy_transforms(demo.m_world, vbo, positionOffset, vboFloat4Capacity,
                                          nullptr);
}

/* createScene(renderer, physicsSim, useConvexHeightfield, fileName)
 * (gpu_rigidbody_pipeline2/BulletDataExtractor.cpp:454-520) with
 * btBulletDataExtractor::convertAllObjects (:594-700): registers the .bullet
 * file's boxes, convex hulls and bodies (static ones only with
 * keepStaticObjects, the reference's global of that name; otherwise its
 * pyramid ground grid), then writeBodiesToGpu.  Returns false when the file
 * cannot be read or is not a float .bullet file (the reference then builds
 * its programmatic scene); last_status() tells which. */
inline bool createSceneFromBulletFile(CLPhysicsDemo &demo, const char *fileName,
                                      bool keepStaticObjects = false) {
    mi_rb_bullet_summary summary;
    demo.m_status = mi_rb_load_bullet(demo.m_world, fileName, keepStaticObjects ? 1 : 0, &summary);
    if (demo.m_status != MI_RB_OK) return false;
    demo.writeBodiesToGpu();
    return demo.m_status == MI_RB_OK;
}

#endif
