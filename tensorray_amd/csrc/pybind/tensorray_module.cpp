// `TensorRay` Python module: the reference's binding surface
// (TensorRay.cpp:48-222) over the host classes of host_scene.h, which in
// turn drive the MI355X kernels through the C ABI of include/trc.h.
#include <pybind11/numpy.h>
#include <pybind11/operators.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "host_scene.h"

namespace py = pybind11;
using namespace py::literals;
using namespace trpy;

PYBIND11_MODULE(TensorRay, m) {
  m.doc() = "Path-space differentiable renderer (MI355X / HIP): TensorRay";
  m.def("env_create", [](int device) { env_create(device); }, "device"_a = 0);
  m.def("env_release", &env_release);
  m.def("get_num_param", &get_num_param);
  m.def("get_size_param", [](py::list sizes) {
    std::vector<int> s;
    int total = get_size_param(s);
    for (int v : s) sizes.append(v);
    return total;
  });
  m.def("get_var", [](int index, FloatPtr p, int offset) { get_var(index, p.p, offset); }, "index"_a, "ptr_data"_a,
        "offset"_a = 0);
  m.def("set_var", [](int index, FloatPtr p, int offset) { set_var(index, p.p, offset); }, "index"_a, "ptr_data"_a,
        "offset"_a = 0);
  m.def("get_grad", [](int index, FloatPtr p, int offset, bool clear) { get_grad(index, p.p, offset, clear); },
        "index"_a, "ptr_data"_a, "offset"_a = 0, "clear_grad"_a = true);
  m.def("set_rnd_seed", [](uint64_t seed) { set_rnd_seed(seed); });
  m.def("to_tensor", [](FloatPtr p, const TensorShape& s) { return Tensorf::from_ptr(p.p, s); });
  m.def("tensor_to_torch", [](const Tensorf& t, FloatPtr p) { t.value_to(p.p); });
  m.def("torch_to_tensor", [](FloatPtr p, Tensorf& t) {
    if (!t.dev) throw std::runtime_error("torch_to_tensor: target tensor has no storage");
    Tensorf src = Tensorf::from_ptr(p.p, t.shape);
    t.dev = src.dev;
  });
  m.def("detach", [](const Tensorf& t) {
    Tensorf r = t;
    r.requires_grad = false;
    return r;
  });

  py::class_<FloatPtr>(m, "float_ptr").def(py::init<std::size_t>());
  py::class_<IntPtr>(m, "int_ptr").def(py::init<std::size_t>());

  py::class_<Vector3f>(m, "Vector3f")
      .def(py::init<float, float, float>())
      .def_readwrite("x", &Vector3f::x)
      .def_readwrite("y", &Vector3f::y)
      .def_readwrite("z", &Vector3f::z);

  py::enum_<VecType>(m, "VecType")
      .value("Scalar1", kScalar1)
      .value("Vec2", kVec2)
      .value("Vec3", kVec3)
      .value("Vec4", kVec4)
      .value("Mat4x4", kMat4x4)
      .export_values();

  py::class_<TensorShape>(m, "TensorShape")
      .def(py::init([](std::vector<int> d, VecType v) { return TensorShape(d, (int)v); }))
      .def(py::init([](std::vector<int> d, int v) { return TensorShape(d, v); }))
      .def("size", &TensorShape::size);

  py::class_<Tensorf>(m, "Tensorf")
      .def(py::init<>())
      // requiresGrad registers the scalar in the parameter pool, as the
      // reference's SetRequiresGrad does (Tensor.h:1675-1685)
      .def(py::init([](float v, bool rg) {
             Tensorf t = Tensorf::scalar(v, rg);
             if (rg) t.param = register_scalar_param(v);
             return t;
           }),
           "val"_a = 0.0f, "requiresGrad"_a = false)
      .def(py::init([](float x, float y, float z, bool rg) { return Tensorf::vec3(x, y, z, rg); }), "x"_a = 0.0f,
           "y"_a = 0.0f, "z"_a = 0.0f, "requiresGrad"_a = false)
      // 4x4 matrix from its entries in row order (the reference's 16-value
      // constructor, TensorRay.cpp:90-98; used for Shape.rotate_matrix)
      .def(py::init([](float m00, float m01, float m02, float m03, float m10, float m11, float m12, float m13,
                       float m20, float m21, float m22, float m23, float m30, float m31, float m32, float m33) {
        Tensorf t;
        t.host = {m00, m01, m02, m03, m10, m11, m12, m13, m20, m21, m22, m23, m30, m31, m32, m33};
        return t;
      }))
      .def(py::init([](FloatPtr p, const TensorShape& s) { return Tensorf::from_ptr(p.p, s); }))
      .def("__mul__", [](const Tensorf& a, const Tensorf& b) { return tensor_mul(a, b); })
      .def("__mul__", [](const Tensorf& a, float f) { return tensor_mul(a, Tensorf::scalar(f, false)); })
      .def("__rmul__", [](const Tensorf& a, float f) { return tensor_mul(Tensorf::scalar(f, false), a); })
      .def("size", &Tensorf::size)
      .def("requires_grad", [](const Tensorf& t) { return t.requires_grad; })
      .def("set_requires_grad", [](Tensorf& t, bool v) { t.requires_grad = v; })
      .def("value_to_torch", [](const Tensorf& t, FloatPtr p) { t.value_to(p.p); })
      .def("grad_to_torch", [](const Tensorf& t, FloatPtr p) {
        // image tensors carry no gradient of their own here (gradients are
        // read per parameter with get_grad); report zeros of the right size
        Tensorf z;
        z.host.assign((size_t)t.size(), 0.0f);
        z.value_to(p.p);
      });

  py::class_<Primitive, std::shared_ptr<Primitive>>(m, "Shape")
      .def_readwrite("obj_center", &Primitive::obj_center)
      .def_readwrite("rotate_matrix", &Primitive::rotate_matrix)
      .def_readwrite("translation", &Primitive::translation)
      .def_readwrite("rotate_axis", &Primitive::rotate_axis)
      .def_readwrite("rotate_angle", &Primitive::rotate_angle)
      .def_readwrite("vertex_id", &Primitive::vertex_id)
      .def_readwrite("vertex_translation", &Primitive::vertex_translation)
      .def("diff_translation", &Primitive::diff_translation)
      .def("diff_rotation", &Primitive::diff_rotation, "rotate_axis"_a)
      .def("diff_all_vertex_pos", &Primitive::diff_all_vertex_pos)
      .def("get_vertex_count", &Primitive::vertex_count)
      .def("get_face_count", &Primitive::face_count)
      .def("get_edge_count", &Primitive::edge_count)
      .def("get_face_indices", [](const Primitive& p, IntPtr out) { p.face_indices(out.p); })
      .def("get_edge_data", [](const Primitive& p, IntPtr out) { p.edge_data(out.p); })
      .def("get_obj_center", &Primitive::get_obj_center)
      .def("export_mesh", &Primitive::export_mesh, "filename"_a);

  py::class_<BSDF, std::shared_ptr<BSDF>>(m, "BSDF")
      .def("diff_texture", &BSDF::diff_texture, "name"_a = std::string());

  py::class_<Light, std::shared_ptr<Light>>(m, "Light");

  py::class_<Camera, std::shared_ptr<Camera>>(m, "Camera")
      .def(py::init<>())
      .def(py::init<const Vector3f&, const Vector3f&, const Vector3f&, int, int, float>())
      .def_readonly("width", &Camera::width)
      .def_readonly("height", &Camera::height)
      .def_property_readonly("posTensor", &Camera::pos_tensor)
      .def("init", &Camera::init)
      .def("update", &Camera::update, "camera"_a)
      .def("resize", &Camera::resize);

  py::class_<Scene>(m, "Scene")
      .def(py::init<>())
      .def("load_file", &Scene::load_file, "filename"_a)
      .def("configure", &Scene::configure)
      .def("get_width", &Scene::get_width, "index"_a = 0)
      .def("get_height", &Scene::get_height, "index"_a = 0)
      .def_property_readonly("num_bsdf", &Scene::num_bsdf)
      .def_readwrite("shapes", &Scene::shapes)
      .def_readwrite("bsdfs", &Scene::bsdfs)
      .def_readwrite("emitters", &Scene::emitters)
      .def_readwrite("cameras", &Scene::cameras)
      // programmatic assembly (the reference builds scenes from XML only)
      .def("add_bsdf",
           [](Scene& s, int type, std::array<float, 3> albedo, std::array<float, 3> spec, float roughness,
              py::object texels) {
             auto b = std::make_shared<BSDF>();
             b->type = type;
             b->albedo = albedo;
             b->spec = spec;
             b->roughness = roughness;
             if (!texels.is_none()) {
               py::array_t<float, py::array::c_style | py::array::forcecast> t(texels);
               if (t.ndim() != 3 || t.shape(2) != 3) throw std::runtime_error("texels must be [h][w][3]");
               b->tex_h = (int)t.shape(0);
               b->tex_w = (int)t.shape(1);
               b->texels.assign(t.data(), t.data() + t.size());
             }
             s.bsdfs.push_back(b);
             return (int)s.bsdfs.size() - 1;
           },
           "type"_a, "albedo"_a, "spec"_a, "roughness"_a, "texels"_a = py::none())
      .def("add_mesh",
           [](Scene& s, py::array_t<float, py::array::c_style | py::array::forcecast> pos,
              py::array_t<uint32_t, py::array::c_style | py::array::forcecast> idx, int bsdf, py::object uv,
              py::object uv_idx, py::object radiance, bool smooth) {
             auto p = std::make_shared<Primitive>();
             p->raw.assign(pos.data(), pos.data() + pos.size());
             p->idx.assign(idx.data(), idx.data() + idx.size());
             p->bsdf = bsdf;
             p->smooth = smooth;
             if (!uv.is_none()) {
               py::array_t<float, py::array::c_style | py::array::forcecast> u(uv);
               py::array_t<uint32_t, py::array::c_style | py::array::forcecast> ui(uv_idx);
               p->uv.assign(u.data(), u.data() + u.size());
               p->uv_idx.assign(ui.data(), ui.data() + ui.size());
             }
             if (!radiance.is_none()) {
               auto r = radiance.cast<std::array<float, 3>>();
               p->emitter = true;
               p->radiance = r;
               auto l = std::make_shared<Light>();
               l->shape = (int)s.shapes.size();
               l->radiance = r;
               s.emitters.push_back(l);
             }
             s.shapes.push_back(p);
             return (int)s.shapes.size() - 1;
           },
           "pos"_a, "idx"_a, "bsdf"_a, "uv"_a = py::none(), "uv_idx"_a = py::none(), "radiance"_a = py::none(),
           "smooth"_a = false)
      .def("add_camera", [](Scene& s, std::shared_ptr<Camera> c) { s.cameras.push_back(c); });

  py::class_<RenderOptions>(m, "RenderOptions")
      .def(py::init<>())
      .def(py::init<int, int, int, int, int, int, int>(), "seed"_a, "max_bounce"_a, "spp"_a, "sppe"_a, "sppse0"_a,
           "sppse1"_a, "sppe0"_a = 0)
      .def_readwrite("seed", &RenderOptions::seed)
      .def_readwrite("max_bounces", &RenderOptions::max_bounces)
      .def_readwrite("spp", &RenderOptions::spp)
      .def_readwrite("spp_batch", &RenderOptions::spp_batch)
      .def_readwrite("sppe", &RenderOptions::sppe)
      .def_readwrite("sppe_batch", &RenderOptions::sppe_batch)
      .def_readwrite("sppse0", &RenderOptions::sppse0)
      .def_readwrite("sppse0_batch", &RenderOptions::sppse0_batch)
      .def_readwrite("sppse1", &RenderOptions::sppse1)
      .def_readwrite("sppse1_batch", &RenderOptions::sppse1_batch)
      .def_readwrite("sppe0", &RenderOptions::sppe0)
      .def_readwrite("sppe0_batch", &RenderOptions::sppe0_batch)
      .def_readwrite("export_deriv", &RenderOptions::export_deriv)
      .def_readwrite("quiet", &RenderOptions::quiet)
      .def_readwrite("checkpoint", &RenderOptions::checkpoint)
      .def_readwrite("g_direct", &RenderOptions::g_direct)
      .def_readwrite("g_direct_depth", &RenderOptions::g_direct_depth)
      .def_readwrite("g_direct_max_size", &RenderOptions::g_direct_max_size)
      .def_readwrite("g_direct_spp", &RenderOptions::g_direct_spp)
      .def_readwrite("g_direct_thold", &RenderOptions::g_direct_thold)
      .def_readwrite("g_eps", &RenderOptions::g_eps)
      .def_readwrite("g_keep", &RenderOptions::g_keep)
      .def_readwrite("g_seed", &RenderOptions::g_seed);

  py::class_<Integrator>(m, "Integrator");
  py::class_<PathTracer, Integrator>(m, "PathTracer")
      .def(py::init<>())
      .def("set_param", &PathTracer::set_param)
      .def("renderC", &PathTracer::render_c)
      .def("renderD", &PathTracer::render_d);
  py::class_<BoundaryIntegrator, Integrator>(m, "BoundaryIntegrator")
      .def("set_param", &BoundaryIntegrator::set_param)
      .def("renderD", &BoundaryIntegrator::render_d);
  py::class_<PrimaryEdgeIntegrator, BoundaryIntegrator>(m, "PrimaryEdgeIntegrator").def(py::init<>());
  py::class_<DirectEdgeIntegrator, BoundaryIntegrator>(m, "DirectEdgeIntegrator").def(py::init<>());
  py::class_<IndirectEdgeIntegrator, BoundaryIntegrator>(m, "IndirectEdgeIntegrator").def(py::init<>());
  py::class_<PixelBoundaryIntegrator, BoundaryIntegrator>(m, "PixelBoundaryIntegrator").def(py::init<>());

  m.attr("MAT_DIFFUSE") = TRC_MAT_DIFFUSE;
  m.attr("MAT_MICROFACET") = TRC_MAT_MICROFACET;
  m.attr("MAT_ROUGHCONDUCTOR") = TRC_MAT_ROUGHCONDUCTOR;
}
