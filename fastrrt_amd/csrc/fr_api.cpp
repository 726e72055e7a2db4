/*
 * fr_api.cpp — host planner objects and the C ABI of libfastrrt.so.
 *
 * FrGraph restates the reference's CudaGraph (include/graph.h:42-286) over
 * one device-resident plan slab; FrFastRRT restates FastRRT
 * (include/fastrrt.h, src/fastrrt.cpp) on top of it.  Every per-cell phase
 * runs on the GPU (fr_kernels.hip); the host keeps only the reference's host
 * logic: the timeout clock, the loop protocol, the text dump and the Hermite
 * resampling of the output path (src/interpolation.cpp).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/fastrrt.h"
#include "fr_engine.h"
#include "fr_math.h"

namespace {

struct Waypoint {
    int x, z;
    float heading;
};

/* ------------------------------------------------------------------------ */
/* Hermite resampling of the output path (src/interpolation.cpp:36-135)     */
/* ------------------------------------------------------------------------ */

int host_trunc(double v)
{
    /* the reference runs this on the host: NaN / overflow -> INT_MIN */
    if (v != v || v >= 2147483648.0 || v <= -2147483649.0)
        return (-2147483647 - 1);
    return (int)v;
}

std::vector<Waypoint> interpolate_hermite_curve(int W, int H, Waypoint p1, Waypoint p2)
{
    std::vector<Waypoint> curve;
    const double ddx = (double)(p2.x - p1.x), ddz = (double)(p2.z - p1.z);
    const int num_points = fr_trunc_d(std::sqrt(ddx * ddx + ddz * ddz));
    if (num_points > 0)
        curve.reserve(num_points);
    const float dx = (float)(p2.x - p1.x), dz = (float)(p2.z - p1.z);
    const float d = sqrtf(dx * dx + dz * dz);
    const float a1 = (float)((double)p1.heading - FR_PI_D / 2);
    const float a2 = (float)((double)p2.heading - FR_PI_D / 2);
    float s1, c1, s2, c2;
    fr_sincosf(a1, &s1, &c1);
    fr_sincosf(a2, &s2, &c2);
    const float t1x = d * c1, t1y = d * s1, t2x = d * c2, t2y = d * s2;
    int last_x = -1, last_z = -1;
    for (int i = 0; i < num_points; ++i) {
        const float t = (float)i / (float)(num_points - 1);
        const float t2 = t * t, t3 = t2 * t;
        const float h00 = 2 * t3 - 3 * t2 + 1;
        const float h10 = t3 - 2 * t2 + t;
        const float h01 = -2 * t3 + 3 * t2;
        const float h11 = t3 - t2;
        const float x = h00 * (float)p1.x + h10 * t1x + h01 * (float)p2.x + h11 * t2x;
        const float z = h00 * (float)p1.z + h10 * t1y + h01 * (float)p2.z + h11 * t2y;
        if (x < 0 || x >= W)
            continue;
        if (z < 0 || z >= H)
            continue;
        const int cx = host_trunc((double)roundf(x));
        const int cz = host_trunc((double)roundf(z));
        if (cx == last_x && cz == last_z)
            continue;
        if (cx < 0 || cx >= W)
            continue;
        if (cz < 0 || cz >= H)
            continue;
        const float t00 = 6 * t2 - 6 * t;
        const float t10 = 3 * t2 - 4 * t + 1;
        const float t01 = -6 * t2 + 6 * t;
        const float t11 = 3 * t2 - 2 * t;
        const float dxx = t00 * (float)p1.x + t10 * t1x + t01 * (float)p2.x + t11 * t2x;
        const float dzz = t00 * (float)p1.z + t10 * t1y + t01 * (float)p2.z + t11 * t2y;
        const float heading = (float)((double)fr_atan2f(dzz, dxx) + FR_HALF_PI_D);
        curve.push_back({cx, cz, heading});
        last_x = cx;
        last_z = cz;
    }
    return curve;
}

std::vector<Waypoint> interpolate_path(const std::vector<Waypoint> &cp, int W, int H)
{
    std::vector<Waypoint> out;
    if (cp.size() < 2)
        return out; /* the reference prints a notice and returns {} */
    Waypoint p1 = cp[0];
    for (size_t i = 1; i < cp.size(); ++i) {
        auto part = interpolate_hermite_curve(W, H, p1, cp[i]);
        out.insert(out.end(), part.begin(), part.end());
        p1 = cp[i];
    }
    return out;
}

float *convert_path(const std::vector<Waypoint> &path)
{
    /* python_bind.cpp:114-136: [n, (x, z, heading) x n] */
    float *res = new float[3 * path.size() + 1];
    res[0] = (float)path.size();
    for (size_t i = 0; i < path.size(); ++i) {
        res[3 * i + 1] = (float)path[i].x;
        res[3 * i + 2] = (float)path[i].z;
        res[3 * i + 3] = path[i].heading;
    }
    return res;
}

} // namespace

/* ------------------------------------------------------------------------ */
/* FrGraph: the reference CudaGraph over one device slab                   */
/* ------------------------------------------------------------------------ */

class FrGraph {
  public:
    FrGraph(int W, int H)
        : eng(1, (long)W * H + 1)
    {
        if (W <= 0 || H <= 0 || (long)W * H > (1L << 28))
            throw std::runtime_error("invalid grid size");
        try {
            slab.allocate(W, H, 0, 0, 0, true, eng.stream());
            prm.W = W;
            prm.H = H;
            /* coordinate origin at the grid centre (graph.cu:258-261,279) */
            prm.og_x = (float)(W / 2);
            prm.og_y = (float)(H / 2);
            FR_CHECK(hipMalloc(&d_tmp, 64 * sizeof(int)));
            FR_CHECK(hipMalloc(&d_region, sizeof(uint32_t) * ((W / 16 + 1) * (H / 16 + 1) + 1)));
            FR_CHECK(hipMemsetAsync(d_tmp, 0, 64 * sizeof(int), eng.stream()));
            rebind();
            std::vector<int> one{0};
            eng.set_active(one);
            eng.reset_ctl(0, FR_IDLE);
        } catch (...) {
            /* a half-built graph runs no destructor */
            (void)hipStreamSynchronize(eng.stream());
            free_buffers();
            throw;
        }
    }
    ~FrGraph()
    {
        eng.sync();
        free_buffers();
    }

    void free_buffers()
    {
        if (d_tmp)
            (void)hipFree(d_tmp);
        if (d_region)
            (void)hipFree(d_region);
        if (d_graph_costs)
            (void)hipFree(d_graph_costs);
        if (d_apf_bits_)
            (void)hipFree(d_apf_bits_);
        if (d_apf_tab_)
            (void)hipFree(d_apf_tab_);
        if (d_fc_)
            (void)hipFree(d_fc_);
        d_fc_ = nullptr;
        fc_cap_ = 0;
        d_tmp = nullptr;
        d_region = nullptr;
        d_graph_costs = nullptr;
        d_apf_bits_ = nullptr;
        d_apf_tab_ = nullptr;
    }

    int W() const { return prm.W; }
    int H() const { return prm.H; }
    int N() const { return prm.W * prm.H; }
    hipStream_t st() { return eng.stream(); }

    /* graph.cu:286-304 */
    void set_physical(float perc_w, float perc_h, double steer_rad, float vehicle_len, float max_curvature)
    {
        perc_w_ = perc_w;
        perc_h_ = perc_h;
        prm.max_steer_rad = steer_rad;
        prm.lr = (double)(vehicle_len / 2.0f);
        const float t = fr_tanf((float)steer_rad);
        double mc;
        if (max_curvature < 0)
            mc = (double)(2.0f * t) / (0.5 * (double)vehicle_len * (double)sqrtf(4.0f + t));
        else
            mc = (double)max_curvature;
        prm.max_curv_loop = (float)mc;
        has_physical = true;
        rebind();
    }

    /* graph.cu:306-314 */
    void set_search(int mdx, int mdz, int lbx, int lbz, int ubx, int ubz)
    {
        prm.min_dx = mdx;
        prm.min_dz = mdz;
        prm.lbx = lbx;
        prm.lbz = lbz;
        prm.ubx = ubx;
        prm.ubz = ubz;
        const int mx = fr_trunc_f((float)mdx / 2.0f), mz = fr_trunc_f((float)mdz / 2.0f);
        if (mx < 0 || mz < 0 || mx > 512 || mz > 512)
            throw std::runtime_error("min distance out of range");
        slab.set_footprint(mx, mz, st());
        frame_prepped = nullptr;
        rebind();
    }

    /* graph.cu:316-338 (the graph's own table, used by derivate_node) */
    void set_class_costs(const float *costs, int n)
    {
        graph_costs.assign(costs, costs + n);
        if (d_graph_costs)
            (void)hipFree(d_graph_costs);
        d_graph_costs = nullptr;
        FR_CHECK(hipMalloc(&d_graph_costs, sizeof(float) * std::max(1, n)));
        if (n)
            FR_CHECK(hipMemcpyAsync(d_graph_costs, graph_costs.data(), sizeof(float) * n, hipMemcpyHostToDevice,
                                    st()));
        eng.sync();
        frame_prepped = nullptr;
    }

    void rebind()
    {
        FrFrameView v;
        v.xyz = bound_xyz;
        v.costs = bound_costs;
        v.n_classes = bound_n;
        eng.bind(0, slab, prm, v);
    }

    /* Prepare the obstacle planes of `frame` read through `costs`. */
    void prep_frame(const FrFrame *frame, bool use_frame_costs)
    {
        if (!frame)
            throw std::runtime_error("null frame");
        if (frame->W != prm.W || frame->H != prm.H)
            throw std::runtime_error("frame size differs from the grid size");
        const float *c = use_frame_costs ? frame->d_costs : d_graph_costs;
        const int n = use_frame_costs ? frame->n_classes : (int)graph_costs.size();
        if (!c)
            throw std::runtime_error("class costs are not set");
        if (frame_prepped == frame && prepped_costs == c && prepped_gen == frame_gen(frame))
            return;
        bound_xyz = frame->d_xyz;
        bound_costs = c;
        bound_n = n;
        rebind();
        run_setup(FR_GOAL_NONE);
        frame_prepped = frame;
        prepped_costs = c;
        prepped_gen = frame_gen(frame);
    }

    /* prep + distance planes, and the goal-connection planes per goal_mode */
    void run_setup(int goal_mode)
    {
        std::vector<int> one{0};
        eng.setup(one, goal_mode);
        if (goal_mode != FR_GOAL_NONE)
            gc_pending = goal_mode == FR_GOAL_LAZY;
    }

    /* evaluate what a lazy goal-connection pass has not evaluated yet */
    void goal_complete()
    {
        if (gc_pending) {
            eng.goal_complete(0);
            gc_pending = false;
        }
    }

    static unsigned long frame_gen(const FrFrame *f) { return f ? gen_of(f) : 0; }
    static unsigned long gen_of(const FrFrame *f);

    /* ---- cell operations ---- */
    bool in(int x, int z) const { return x >= 0 && x < prm.W && z >= 0 && z < prm.H; }

    void add(int x, int z, float h, int px, int pz, float cost)
    {
        fr_launch_graph_add(eng.launch(), 0, x, z, h, px, pz, cost, FR_TYPE_NODE, 1, st());
        dirty();
    }
    void add_temporary(int x, int z, float h, int px, int pz, float cost)
    {
        fr_launch_graph_add(eng.launch(), 0, x, z, h, px, pz, cost, FR_TYPE_TEMP, 0, st());
        dirty();
    }
    void remove(int x, int z)
    {
        if (!in(x, z))
            return;
        uint8_t zero = 0;
        FR_CHECK(hipMemcpyAsync(slab.type + (z * prm.W + x), &zero, 1, hipMemcpyHostToDevice, st()));
        eng.sync();
        dirty();
    }
    void set_collision(int x, int z, int npx, int npz, float h, float cost)
    {
        fr_launch_set_collision(eng.launch(), 0, x, z, npx, npz, h, cost, st());
        node_collision = true;
        dirty();
    }
    /* graph.cu:416-427 */
    void clear()
    {
        fr_launch_clear(eng.launch(), 0, N(), st());
        FrCtl c;
        eng.read_ctl(0, &c);
        c.goal_reached = 0;
        eng.write_ctl(0, c);
        node_collision = false;
        dirty();
    }

    template <typename T> T read1(const T *p, int x, int z)
    {
        T v;
        FR_CHECK(hipMemcpyAsync(&v, p + (z * prm.W + x), sizeof(T), hipMemcpyDeviceToHost, st()));
        eng.sync();
        return v;
    }

    int get_type(int x, int z)
    {
        if (!in(x, z))
            return -1;
        return read1(slab.type, x, z);
    }
    bool check_in_graph(int x, int z) { return in(x, z) && read1(slab.type, x, z) == FR_TYPE_NODE; }
    std::pair<int, int> get_parent(int x, int z)
    {
        if (!in(x, z) || read1(slab.type, x, z) == FR_TYPE_NULL)
            return {-1, -1};
        return {read1(slab.px, x, z), read1(slab.pz, x, z)};
    }
    float get_heading(int x, int z)
    {
        /* graph.cu:455-459 reads without a bounds check; out of range -> 0 here */
        if (!in(x, z))
            return 0.0f;
        return read1(slab.heading, x, z);
    }
    float get_cost(int x, int z) { return in(x, z) ? read1(slab.cost, x, z) : -1.0f; }
    int child_count(int x, int z) { return in(x, z) ? read1(slab.derive, x, z) : 0; }

    std::vector<int> count_types()
    {
        FR_CHECK(hipMemsetAsync(d_tmp, 0, 8 * sizeof(int), st()));
        fr_launch_count(eng.launch(), 0, N(), d_tmp, st());
        std::vector<int> h(8);
        FR_CHECK(hipMemcpyAsync(h.data(), d_tmp, 8 * sizeof(int), hipMemcpyDeviceToHost, st()));
        eng.sync();
        return h;
    }
    int count(int type)
    {
        auto c = count_types();
        return (type >= 0 && type < 6) ? c[type] : 0;
    }
    int count_all()
    {
        auto c = count_types();
        return c[1] + c[2] + c[3] + c[4] + c[5];
    }

    void accept_derived() { fr_launch_accept(eng.launch(), 0, N(), st()); eng.sync(); dirty(); }

    /* graph_collision.cu:106-126 */
    void solve_collisions()
    {
        const int nodes = count(FR_TYPE_NODE);
        if (nodes <= 2)
            return;
        fr_launch_solve_collisions(eng.launch(), 0, N(), nodes, st());
        eng.sync();
        node_collision = false;
        dirty();
    }

    /* graph.cu:224-244 -> expand_node */
    std::pair<int, int> derivate(const FrFrame *frame, float steer, float path_size, float v, int x, int z)
    {
        if (!has_physical)
            throw std::runtime_error("physical parameters are not set");
        prep_frame(frame, false);
        FR_CHECK(hipMemsetAsync(d_tmp, 0, 4 * sizeof(int), st()));
        fr_launch_derivate(eng.launch(), 0, steer, path_size, v, x, z, d_tmp, st());
        int h[4];
        FR_CHECK(hipMemcpyAsync(h, d_tmp, 4 * sizeof(int), hipMemcpyDeviceToHost, st()));
        eng.sync();
        if (h[2])
            node_collision = true;
        dirty();
        return {h[0], h[1]};
    }

    /* CudaGraph::checkFeasibleConnection (kinematics.cu:308-326) for n
     * queries (start x, start z, end x, end z): the graph's class costs,
     * footprint, ego box and headings, as there.  Outputs per query: the
     * verdict, and check_kinematic_connection_start_end's final heading and
     * path cost (either may be null). */
    void feasible_connections(const FrFrame *frame, const int32_t *q, int n, int velocity, uint8_t *ok, double *fh,
                              float *pc)
    {
        if (!has_physical)
            throw std::runtime_error("physical parameters are not set");
        if (n < 0 || (n > 0 && (!q || !ok)))
            throw std::runtime_error("check_feasible_connections: invalid query or output buffer");
        if (n == 0)
            return;
        prep_frame(frame, false);
        const size_t need = 32 * (size_t)n; /* int4 query + double + float + byte, 16/8/4/1 (+3 spare) */
        if (need > fc_cap_) {
            eng.sync();
            if (d_fc_)
                (void)hipFree(d_fc_);
            d_fc_ = nullptr;
            fc_cap_ = 0;
            FR_CHECK(hipMalloc(&d_fc_, need));
            fc_cap_ = need;
        }
        int4 *d_q = reinterpret_cast<int4 *>(d_fc_);
        double *d_h = reinterpret_cast<double *>(d_fc_ + 16 * (size_t)n);
        float *d_c = reinterpret_cast<float *>(d_fc_ + 24 * (size_t)n);
        uint8_t *d_ok = d_fc_ + 28 * (size_t)n;
        FR_CHECK(hipMemcpyAsync(d_q, q, 16 * (size_t)n, hipMemcpyHostToDevice, st()));
        fr_launch_feasible_connections(eng.launch(), 0, d_q, n, velocity, d_ok, d_h, d_c, st());
        FR_CHECK(hipGetLastError());
        FR_CHECK(hipMemcpyAsync(ok, d_ok, (size_t)n, hipMemcpyDeviceToHost, st()));
        if (fh)
            FR_CHECK(hipMemcpyAsync(fh, d_h, 8 * (size_t)n, hipMemcpyDeviceToHost, st()));
        if (pc)
            FR_CHECK(hipMemcpyAsync(pc, d_c, 4 * (size_t)n, hipMemcpyDeviceToHost, st()));
        eng.sync();
    }

    /* graph_expansion.cu:246-261 (the frame's class costs; the footprint and
     * ego box are the graph's, as the frame handle here carries no search
     * parameters).  first: the goal pixel is an obstacle; second: the
     * curve's result (cost, -1, -2). */
    std::pair<bool, float> can_connect_hermite(const FrFrame *frame, int x, int z, int gx, int gz, int goal_heading)
    {
        if (!has_physical)
            throw std::runtime_error("physical parameters are not set");
        prep_frame(frame, true);
        fr_launch_hermite_probe(eng.launch(), 0, x, z, gx, gz, goal_heading, (float)prm.max_steer_rad,
                                (float *)d_tmp, st());
        float h[2];
        FR_CHECK(hipMemcpyAsync(h, d_tmp, 2 * sizeof(float), hipMemcpyDeviceToHost, st()));
        eng.sync();
        return {h[1] != 0.0f, h[0]};
    }

    /* graph_direct_goal_connection.cu:200-227 (planes from the frame's costs).
     * lazy: a cell's planes are evaluated when a node's connection window
     * first covers it (same values: the evaluation is a function of the
     * frame, the goal and the cost plane at this call, which is kept). */
    void process_direct_goal_connection(const FrFrame *frame, int gx, int gz, float gh, float max_curv,
                                        bool lazy = false)
    {
        if (!frame)
            throw std::runtime_error("null frame");
        if (max_curv < 0)
            max_curv = prm.max_curv_loop;
        prm.gx = gx;
        prm.gz = gz;
        prm.gh = gh;
        prm.max_curv_goal = max_curv;
        prm.gc_lazy = lazy ? 1 : 0;
        bound_xyz = frame->d_xyz;
        bound_costs = frame->d_costs;
        bound_n = frame->n_classes;
        rebind();
        run_setup(lazy ? FR_GOAL_LAZY : FR_GOAL_EAGER);
        frame_prepped = frame;
        prepped_costs = frame->d_costs;
        prepped_gen = frame_gen(frame);
    }

    /* graph_best_node.cu:224-251 */
    bool check_goal_reached(int gx, int gz, float gh, float tol, float herr)
    {
        FrCtl c;
        eng.read_ctl(0, &c);
        if (c.goal_reached)
            return true;
        fr_launch_goal_scan(eng.launch(), 0, N(), gx, gz, gh, tol, herr, st());
        eng.read_ctl(0, &c);
        return c.goal_reached != 0;
    }

    /* graph_best_node.cu:134-178 */
    std::pair<int, int> find_best_node(const FrFrame *frame, float heading, float radius, int x, int z, float herr,
                                       bool frame_costs)
    {
        if (frame)
            prep_frame(frame, frame_costs);
        FrBestNodeArgs a{x, z, heading, radius, herr};
        fr_launch_best_node(eng.launch(), eng.d_active(), 1, N(), &a, st());
        FrCtl c;
        eng.read_ctl(0, &c);
        if ((unsigned long long)c.bn_cost >= (unsigned long long)(FR_BIG_COST + (1LL << 62)) ||
            c.bn_pos == 0x7FFFFFFF)
            return {-1, -1};
        return {c.bn_pos % prm.W, c.bn_pos / prm.W};
    }

    /* src/graph_check_dag_consistency.cpp:34-76, on a host copy */
    bool check_consistent()
    {
        const int n = N();
        std::vector<uint8_t> ty(n);
        std::vector<int32_t> px(n), pz(n);
        FR_CHECK(hipMemcpyAsync(ty.data(), slab.type, n, hipMemcpyDeviceToHost, st()));
        FR_CHECK(hipMemcpyAsync(px.data(), slab.px, 4 * (size_t)n, hipMemcpyDeviceToHost, st()));
        FR_CHECK(hipMemcpyAsync(pz.data(), slab.pz, 4 * (size_t)n, hipMemcpyDeviceToHost, st()));
        eng.sync();
        int nodes = 0;
        for (int i = 0; i < n; ++i)
            nodes += ty[i] == FR_TYPE_NODE;
        const int max_loop = nodes + 4;
        auto parent = [&](int x, int z, int &ox, int &oz) {
            if (!in(x, z) || ty[z * prm.W + x] == FR_TYPE_NULL) {
                ox = -1;
                oz = -1;
                return;
            }
            ox = px[z * prm.W + x];
            oz = pz[z * prm.W + x];
        };
        for (int pos = 0; pos < n; ++pos) {
            if (ty[pos] == FR_TYPE_NULL)
                continue;
            int x = pos % prm.W, z = pos / prm.W;
            int i = max_loop;
            while (i-- > 0) {
                int qx, qz;
                parent(x, z, qx, qz);
                if (qx == x && qz == z)
                    return false;
                if (qx == -1)
                    break;
                const int pt = in(qx, qz) ? ty[qz * prm.W + qx] : -1;
                if (pt != FR_TYPE_NODE && pt != FR_TYPE_COLLISION && pt != FR_TYPE_CONNECT_TO_GOAL)
                    return false;
                x = qx;
                z = qz;
            }
            if (i <= 0)
                return false;
        }
        return true;
    }

    /* graph_list.cu:139-209: non-NULL cells in raster order */
    struct Listed {
        int x, z, type;
    };
    std::vector<Listed> list_all()
    {
        const int n = N();
        std::vector<uint8_t> ty(n);
        FR_CHECK(hipMemcpyAsync(ty.data(), slab.type, n, hipMemcpyDeviceToHost, st()));
        eng.sync();
        std::vector<Listed> out;
        for (int i = 0; i < n; ++i)
            if (ty[i] != FR_TYPE_NULL)
                out.push_back({i % prm.W, i / prm.W, (int)ty[i]});
        return out;
    }

    /* graph.cu:496-520: "%d %d %d %d %f %f %f %f" per cell */
    void dump(const char *path)
    {
        const int n = N();
        std::vector<uint8_t> ty(n);
        std::vector<int32_t> px(n), pz(n), dv(n);
        std::vector<float> hd(n), co(n), in_(n), di(n);
        FR_CHECK(hipMemcpyAsync(ty.data(), slab.type, n, hipMemcpyDeviceToHost, st()));
        FR_CHECK(hipMemcpyAsync(px.data(), slab.px, 4 * (size_t)n, hipMemcpyDeviceToHost, st()));
        FR_CHECK(hipMemcpyAsync(pz.data(), slab.pz, 4 * (size_t)n, hipMemcpyDeviceToHost, st()));
        FR_CHECK(hipMemcpyAsync(dv.data(), slab.derive, 4 * (size_t)n, hipMemcpyDeviceToHost, st()));
        FR_CHECK(hipMemcpyAsync(hd.data(), slab.heading, 4 * (size_t)n, hipMemcpyDeviceToHost, st()));
        FR_CHECK(hipMemcpyAsync(co.data(), slab.cost, 4 * (size_t)n, hipMemcpyDeviceToHost, st()));
        FR_CHECK(hipMemcpyAsync(in_.data(), slab.intrinsic, 4 * (size_t)n, hipMemcpyDeviceToHost, st()));
        FR_CHECK(hipMemcpyAsync(di.data(), slab.direct, 4 * (size_t)n, hipMemcpyDeviceToHost, st()));
        eng.sync();
        FILE *fp = fopen(path, "w");
        if (!fp)
            throw std::runtime_error(std::string("cannot open ") + path);
        for (int i = 0; i < n; ++i)
            fprintf(fp, "%d %d %d %d %f %f %f %f\n", px[i], pz[i], (int)ty[i], dv[i], hd[i], co[i], in_[i], di[i]);
        fclose(fp);
    }

    /* graph.cu:522-546 */
    void read_dump(const char *path)
    {
        FILE *fp = fopen(path, "r");
        if (!fp)
            throw std::runtime_error(std::string("cannot open ") + path);
        const int n = N();
        std::vector<uint8_t> ty(n);
        std::vector<int32_t> px(n), pz(n), dv(n);
        std::vector<float> hd(n), co(n), in_(n), di(n);
        FR_CHECK(hipMemcpyAsync(ty.data(), slab.type, n, hipMemcpyDeviceToHost, st()));
        FR_CHECK(hipMemcpyAsync(px.data(), slab.px, 4 * (size_t)n, hipMemcpyDeviceToHost, st()));
        FR_CHECK(hipMemcpyAsync(pz.data(), slab.pz, 4 * (size_t)n, hipMemcpyDeviceToHost, st()));
        FR_CHECK(hipMemcpyAsync(dv.data(), slab.derive, 4 * (size_t)n, hipMemcpyDeviceToHost, st()));
        FR_CHECK(hipMemcpyAsync(hd.data(), slab.heading, 4 * (size_t)n, hipMemcpyDeviceToHost, st()));
        FR_CHECK(hipMemcpyAsync(co.data(), slab.cost, 4 * (size_t)n, hipMemcpyDeviceToHost, st()));
        FR_CHECK(hipMemcpyAsync(in_.data(), slab.intrinsic, 4 * (size_t)n, hipMemcpyDeviceToHost, st()));
        FR_CHECK(hipMemcpyAsync(di.data(), slab.direct, 4 * (size_t)n, hipMemcpyDeviceToHost, st()));
        eng.sync();
        for (int i = 0; i < n; ++i) {
            int t;
            if (fscanf(fp, "%d %d %d %d %f %f %f %f\n", &px[i], &pz[i], &t, &dv[i], &hd[i], &co[i], &in_[i],
                       &di[i]) != 8)
                break; /* the reference keeps whatever was read */
            if (t < 0 || t > 255)
                t = FR_TYPE_NULL; /* node types are 0..5 */
            ty[i] = (uint8_t)t;
        }
        fclose(fp);
        FR_CHECK(hipMemcpyAsync(slab.type, ty.data(), n, hipMemcpyHostToDevice, st()));
        FR_CHECK(hipMemcpyAsync(slab.px, px.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st()));
        FR_CHECK(hipMemcpyAsync(slab.pz, pz.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st()));
        FR_CHECK(hipMemcpyAsync(slab.derive, dv.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st()));
        FR_CHECK(hipMemcpyAsync(slab.heading, hd.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st()));
        FR_CHECK(hipMemcpyAsync(slab.cost, co.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st()));
        FR_CHECK(hipMemcpyAsync(slab.intrinsic, in_.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st()));
        FR_CHECK(hipMemcpyAsync(slab.direct, di.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st()));
        eng.sync();
        dirty();
        FrCtl c;
        eng.read_ctl(0, &c);
        c.full_check = 1;
        eng.write_ctl(0, c);
    }

    /* binary copy of the grid planes (any pointer may be null) */
    void export_state(int32_t *opx, int32_t *opz, int32_t *otype, int32_t *oderive, float *oheading, float *ocost,
                      uint8_t *ogc_flag, float *ogc_heading, float *ogc_cost)
    {
        const size_t n = (size_t)N();
        std::vector<uint8_t> ty;
        if (ogc_flag || ogc_heading || ogc_cost)
            goal_complete();
        if (opx)
            FR_CHECK(hipMemcpyAsync(opx, slab.px, 4 * n, hipMemcpyDeviceToHost, st()));
        if (opz)
            FR_CHECK(hipMemcpyAsync(opz, slab.pz, 4 * n, hipMemcpyDeviceToHost, st()));
        if (oderive)
            FR_CHECK(hipMemcpyAsync(oderive, slab.derive, 4 * n, hipMemcpyDeviceToHost, st()));
        if (oheading)
            FR_CHECK(hipMemcpyAsync(oheading, slab.heading, 4 * n, hipMemcpyDeviceToHost, st()));
        if (ocost)
            FR_CHECK(hipMemcpyAsync(ocost, slab.cost, 4 * n, hipMemcpyDeviceToHost, st()));
        if (ogc_flag)
            FR_CHECK(hipMemcpyAsync(ogc_flag, slab.gc_flag, n, hipMemcpyDeviceToHost, st()));
        if (ogc_heading)
            FR_CHECK(hipMemcpyAsync(ogc_heading, slab.gc_heading, 4 * n, hipMemcpyDeviceToHost, st()));
        if (ogc_cost)
            FR_CHECK(hipMemcpyAsync(ogc_cost, slab.gc_cost, 4 * n, hipMemcpyDeviceToHost, st()));
        if (otype) {
            ty.resize(n);
            FR_CHECK(hipMemcpyAsync(ty.data(), slab.type, n, hipMemcpyDeviceToHost, st()));
        }
        eng.sync();
        if (otype)
            for (size_t i = 0; i < n; ++i)
                otype[i] = ty[i];
    }

    std::vector<float> intrinsic_costs()
    {
        std::vector<float> h(N());
        FR_CHECK(hipMemcpyAsync(h.data(), slab.intrinsic, 4 * (size_t)N(), hipMemcpyDeviceToHost, st()));
        eng.sync();
        return h;
    }

    /* apf.cu:123-155.  The graph's own class table and ego box are used, as
     * in the reference (_classCosts, _searchSpaceParams). */
    FrApfArgs apf_args(const FrFrame *frame)
    {
        if (!frame)
            throw std::runtime_error("null frame");
        if (frame->W != prm.W || frame->H != prm.H)
            throw std::runtime_error("frame size differs from the grid size");
        if (!d_graph_costs)
            throw std::runtime_error("class costs are not set");
        FrApfArgs a;
        a.W = prm.W;
        a.H = prm.H;
        a.xyz = frame->d_xyz;
        a.costs = d_graph_costs;
        a.n_classes = (int)graph_costs.size();
        a.lbx = prm.lbx;
        a.lbz = prm.lbz;
        a.ubx = prm.ubx;
        a.ubz = prm.ubz;
        a.intrinsic = slab.intrinsic;
        return a;
    }

    void apf_repulsion(const FrFrame *frame, float kr, int radius)
    {
        const FrApfArgs a = apf_args(frame);
        const int w = fr_apf_window(radius, prm.W, prm.H);
        if (w > FR_APF_MAX_WINDOW)
            throw std::runtime_error("repulsion window beyond 2048 px (radius and grid both larger)");
        const size_t words = (size_t)prm.H * ((prm.W + 63) / 64);
        const size_t tab = (size_t)(w + 1) * (w + 1);
        if (words > apf_words_ || tab > apf_tab_) {
            eng.sync();
            if (d_apf_bits_)
                (void)hipFree(d_apf_bits_);
            if (d_apf_tab_)
                (void)hipFree(d_apf_tab_);
            d_apf_bits_ = nullptr;
            d_apf_tab_ = nullptr;
            apf_words_ = apf_tab_ = 0;
            FR_CHECK(hipMalloc(&d_apf_bits_, 8 * words));
            FR_CHECK(hipMalloc(&d_apf_tab_, 4 * tab + 4 * (size_t)(FR_APF_MAX_WINDOW + 1)));
            apf_words_ = words;
            apf_tab_ = tab;
        }
        fr_launch_apf_repulsion(a, (float)(kr * 0.5), radius, d_apf_bits_, d_apf_tab_, (int *)(d_apf_tab_ + tab),
                                st());
        FR_CHECK(hipGetLastError());
        eng.sync();
        dirty();
    }

    /* the repulsion above with its scratch in the library's static device
     * array (fr_apf.hip): no hipMalloc/hipFree on any call */
    void apf_repulsion_noalloc(const FrFrame *frame, float kr, int radius)
    {
        const FrApfArgs a = apf_args(frame);
        if (fr_apf_window(radius, prm.W, prm.H) > FR_APF_MAX_WINDOW)
            throw std::runtime_error("repulsion window beyond 2048 px (radius and grid both larger)");
        fr_launch_apf_repulsion_noalloc(a, (float)(kr * 0.5), radius, st());
        dirty();
    }

    void apf_attraction(const FrFrame *frame, float ka, int gx, int gz)
    {
        const FrApfArgs a = apf_args(frame);
        fr_launch_apf_attraction(a, (float)(ka * 0.5), gx, gz, st());
        FR_CHECK(hipGetLastError());
        eng.sync();
        dirty();
    }

    bool gc_flag(int x, int z) { return in(x, z) && read1(slab.gc_flag, x, z) != 0; }
    float gc_cost(int x, int z) { return in(x, z) ? read1(slab.gc_cost, x, z) : 0.0f; }
    float gc_heading(int x, int z) { return in(x, z) ? read1(slab.gc_heading, x, z) : 0.0f; }

    void compute_region_density()
    {
        const int nr = (prm.W / prm.region_side + 1) * (prm.H / prm.region_side + 1);
        FR_CHECK(hipMemsetAsync(d_region, 0, sizeof(uint32_t) * nr, st()));
        fr_launch_region_count(eng.launch(), 0, N(), d_region, st());
        eng.sync();
    }

    void dirty() { goal_valid = false; }

    FrEngine eng;
    FrSlab slab;
    FrPlanParams prm;
    bool has_physical = false;
    bool node_collision = false;
    bool goal_valid = false;
    float perc_w_ = 0, perc_h_ = 0;
    std::vector<float> graph_costs;
    float *d_graph_costs = nullptr;
    const float *bound_xyz = nullptr;
    const float *bound_costs = nullptr;
    int bound_n = 0;
    const FrFrame *frame_prepped = nullptr;
    const float *prepped_costs = nullptr;
    unsigned long prepped_gen = 0;
    bool gc_pending = false; /* lazy goal-connection planes not fully evaluated */
    int *d_tmp = nullptr;
    uint32_t *d_region = nullptr;
    /* potential-field scratch: edge bit plane, offset table + disc spans */
    uint64_t *d_apf_bits_ = nullptr;
    float *d_apf_tab_ = nullptr;
    size_t apf_words_ = 0, apf_tab_ = 0;
    /* feasible-connection queries and results (grown on demand) */
    uint8_t *d_fc_ = nullptr;
    size_t fc_cap_ = 0;
};

/* frames carry a generation counter so a re-uploaded frame is re-prepared */
struct FrFrameObj : FrFrame {
    unsigned long gen = 1;
};
unsigned long FrGraph::gen_of(const FrFrame *f) { return static_cast<const FrFrameObj *>(f)->gen; }

/* ------------------------------------------------------------------------ */
/* FrFastRRT: the reference FastRRT (src/fastrrt.cpp)                       */
/* ------------------------------------------------------------------------ */

class FrFastRRT {
  public:
    FrFastRRT(int W, int H, float perc_w, float perc_h, float steer_deg, float vehicle_len, int lbx, int lbz,
              int ubx, int ubz, float max_curvature)
        : g(W, H)
    {
        lbx_ = lbx;
        lbz_ = lbz;
        ubx_ = ubx;
        ubz_ = ubz;
        herr_tol = (float)(10.0 * FR_TO_RAD_D); /* fastrrt.cpp:39 */
        g.set_physical(perc_w, perc_h, (double)steer_deg * FR_TO_RAD_D, vehicle_len, max_curvature);
    }

    /* fastrrt.cpp:64-81 */
    void set_plan_data(const FrFrame *frame, int sx, int sz, float sh, int gx, int gz, float gh, float v, int mdx,
                       int mdz, int timeout, float max_path, float dist_tol, float herr)
    {
        if (!frame)
            throw std::runtime_error("null frame");
        if (frame->W != g.W() || frame->H != g.H())
            throw std::runtime_error("frame size differs from the planner's grid size");
        has_plan = true;
        frame_ = frame;
        g.prm.sx = sx;
        g.prm.sz = sz;
        g.prm.sh = sh;
        g.prm.v = v;
        timeout_ms = timeout;
        g.prm.max_path = max_path;
        g.prm.dist_tol = dist_tol;
        herr_tol = herr;
        g.prm.herr_tol = herr;
        g.set_search(mdx, mdz, lbx_, lbz_, ubx_, ubz_);
        g.graph_costs = frame->h_costs;
        g.process_direct_goal_connection(frame, gx, gz, gh, 0.8f, true);
    }

    /* fastrrt.cpp:85-98 */
    void search_init()
    {
        if (!has_plan)
            throw std::runtime_error("unable to initialize planning without planning data");
        t0 = std::chrono::steady_clock::now();
        fr_launch_clear(g.eng.launch(), 0, g.N(), g.st());
        g.prm.seed = seed;
        g.prm.stream = stream;
        stream += 1;
        g.rebind();
        fr_launch_graph_add(g.eng.launch(), 0, g.prm.sx, g.prm.sz, g.prm.sh, -1, -1, 0.0f, FR_TYPE_NODE, 1, g.st());
        FR_CHECK(hipMemsetAsync(g.slab.claim, 0xFF, 8 * (size_t)g.N(), g.st()));
        g.eng.reset_ctl(0, FR_SEARCHING);
        ctl_status = FR_SEARCHING;
        last_count = 0;
        iterations = 0;
        g.goal_valid = false;
        g.node_collision = false;
    }

    bool timed_out() const
    {
        if (timeout_ms <= 0)
            return false;
        auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(std::chrono::steady_clock::now() - t0);
        return ms.count() > timeout_ms;
    }

    /* fastrrt.cpp:108-181 */
    bool loop(bool smart)
    {
        if (!has_plan)
            throw std::runtime_error("loop() without planning data");
        if (timed_out())
            return false;
        if ((int)smart != g.prm.smart) {
            g.prm.smart = smart ? 1 : 0;
            g.rebind();
        }
        if (ctl_status != FR_SEARCHING) {
            FrCtl c;
            g.eng.read_ctl(0, &c);
            c.status = FR_SEARCHING;
            g.eng.write_ctl(0, c);
        }
        g.eng.iteration();
        FrCtl c;
        g.eng.read_ctl(0, &c);
        ctl_status = c.status;
        last_count = c.last_temp;
        last_expansions = c.stat_expansions;
        last_new = c.last_temp;
        last_fresh = c.stat_fresh;
        iterations = c.iter;
        g.goal_valid = true;
        goal_flag = c.goal_reached != 0;
        if (c.stat_reseeded)
            return true;
        return !goal_flag;
    }

    /* fastrrt.cpp:198-202 */
    bool goal_reached()
    {
        if (g.goal_valid)
            return goal_flag;
        goal_flag = g.check_goal_reached(g.prm.gx, g.prm.gz, g.prm.gh, g.prm.dist_tol, herr_tol);
        g.goal_valid = true;
        return goal_flag;
    }

    /* fastrrt.cpp:204-232: best node, parent walk (device), copy */
    bool planned_path_device()
    {
        if (!has_plan)
            return false;
        if (!goal_reached())
            return false;
        const FrLaunch &L = g.eng.launch();
        fr_launch_best_node(L, g.eng.d_active(), 1, g.N(), nullptr, g.st());
        fr_launch_walk(L, g.eng.d_active(), 1, g.st());
        return true;
    }

    std::vector<Waypoint> get_planned_path()
    {
        std::vector<Waypoint> res;
        if (!planned_path_device())
            return res;
        FrCtl c;
        g.eng.read_ctl(0, &c);
        if (c.path_len <= 0)
            return res;
        std::vector<float4> p(c.path_len);
        FR_CHECK(hipMemcpyAsync(p.data(), g.slab.path, sizeof(float4) * c.path_len, hipMemcpyDeviceToHost, g.st()));
        g.eng.sync();
        res.reserve(c.path_len);
        for (auto &w : p)
            res.push_back({(int)w.x, (int)w.y, w.z});
        return res;
    }

    /* fastrrt.cpp:183-196 -> graph_optimization.cu:92-195 (pairs on the GPU) */
    bool path_optimize()
    {
        if (timed_out())
            return false;
        if (!planned_path_device())
            return false;
        FrCtl c;
        g.eng.read_ctl(0, &c);
        if (c.path_len < 3)
            return false;
        fr_launch_opt_round(g.eng.launch(), g.eng.d_active(), 1, c.path_len, g.prm.W * g.prm.H, g.st());
        g.eng.read_ctl(0, &c);
        if (c.opt_changed) {
            g.goal_valid = false;
            goal_flag = false;
        }
        return c.opt_changed != 0;
    }

    std::vector<Waypoint> ideal_curve(int gx, int gz, float gh)
    {
        /* fastrrt.cpp:270-281: from the coordinate origin */
        Waypoint s{(int)g.prm.og_x, (int)g.prm.og_y, 0.0f};
        return interpolate_hermite_curve(g.W(), g.H(), s, {gx, gz, gh});
    }

    FrGraph g;
    const FrFrame *frame_ = nullptr;
    bool has_plan = false;
    int timeout_ms = 0;
    float herr_tol = 0;
    std::chrono::steady_clock::time_point t0;
    int last_count = 0;
    int last_expansions = 0, last_new = 0, last_fresh = 0;
    int iterations = 0;
    int ctl_status = FR_IDLE;
    bool goal_flag = false;
    uint64_t seed = 0x5EEDF457ULL, stream = 0;
    int lbx_, lbz_, ubx_, ubz_;
};

/* ------------------------------------------------------------------------ */
/* C ABI                                                                    */
/* ------------------------------------------------------------------------ */

#define FR_TRY(ret_on_error, body)                                                                        \
    try {                                                                                                  \
        fr_set_error("");                                                                                  \
        body                                                                                               \
    } catch (const std::exception &e) {                                                                    \
        fr_set_error(e.what());                                                                            \
        return ret_on_error;                                                                               \
    } catch (...) {                                                                                        \
        fr_set_error("unknown error");                                                                     \
        return ret_on_error;                                                                               \
    }

#define FR_TRY_VOID(body)                                                                                 \
    try {                                                                                                  \
        fr_set_error("");                                                                                  \
        body                                                                                               \
    } catch (const std::exception &e) {                                                                    \
        fr_set_error(e.what());                                                                            \
    } catch (...) {                                                                                        \
        fr_set_error("unknown error");                                                                     \
    }

static FrFastRRT *R(void *p)
{
    if (!p)
        throw std::runtime_error("null FastRRT handle");
    return static_cast<FrFastRRT *>(p);
}
static FrGraph *G(void *p)
{
    if (!p)
        throw std::runtime_error("null graph handle");
    return static_cast<FrGraph *>(p);
}
static const FrFrame *F(void *p)
{
    if (!p)
        throw std::runtime_error("null frame handle");
    return static_cast<const FrFrame *>(p);
}

extern "C" {

const char *fastrrt_last_error(void) { return fr_get_error(); }

void *fastrrt_frame_create(int width, int height, const float *xyz3, const float *class_costs, int n_classes)
{
    FR_TRY(nullptr, {
        if (width <= 0 || height <= 0)
            throw std::runtime_error("invalid frame size");
        if (!xyz3)
            throw std::runtime_error("null pixel buffer");
        if (n_classes < 0 || (n_classes > 0 && !class_costs))
            throw std::runtime_error("invalid class-cost table");
        std::unique_ptr<FrFrameObj> f(new FrFrameObj());
        f->W = width;
        f->H = height;
        f->n_classes = n_classes;
        f->h_costs.assign(class_costs, class_costs + n_classes);
        const size_t bytes = sizeof(float) * 3 * (size_t)width * height;
        FR_CHECK(hipMalloc(&f->d_xyz, bytes));
        FR_CHECK(hipMalloc(&f->d_costs, sizeof(float) * std::max(1, n_classes)));
        FR_CHECK(hipMemcpy(f->d_xyz, xyz3, bytes, hipMemcpyHostToDevice));
        if (n_classes)
            FR_CHECK(hipMemcpy(f->d_costs, class_costs, sizeof(float) * n_classes, hipMemcpyHostToDevice));
        return (void *)f.release();
    })
}

void fastrrt_frame_destroy(void *frame)
{
    FR_TRY_VOID({
        FR_CHECK(hipDeviceSynchronize());
        delete static_cast<FrFrameObj *>(frame);
    })
}

bool fastrrt_frame_update(void *frame, const float *xyz3)
{
    FR_TRY(false, {
        FrFrameObj *f = static_cast<FrFrameObj *>(frame);
        if (!f || !xyz3)
            throw std::runtime_error("null frame or pixel buffer");
        FR_CHECK(hipDeviceSynchronize());
        FR_CHECK(hipMemcpy(f->d_xyz, xyz3, sizeof(float) * 3 * (size_t)f->W * f->H, hipMemcpyHostToDevice));
        f->gen += 1;
        return true;
    })
}

bool fastrrt_frame_read(void *frame, float *xyz3)
{
    FR_TRY(false, {
        FrFrameObj *f = static_cast<FrFrameObj *>(frame);
        if (!f || !xyz3)
            throw std::runtime_error("null frame or pixel buffer");
        FR_CHECK(hipDeviceSynchronize());
        FR_CHECK(hipMemcpy(xyz3, f->d_xyz, sizeof(float) * 3 * (size_t)f->W * f->H, hipMemcpyDeviceToHost));
        return true;
    })
}

bool fastrrt_frame_process_distance_to_goal(void *frame, int goal_x, int goal_z)
{
    FR_TRY(false, {
        FrFrameObj *f = static_cast<FrFrameObj *>(frame);
        if (!f)
            throw std::runtime_error("null frame");
        FR_CHECK(hipDeviceSynchronize());
        fr_launch_frame_distance_to_goal(f->d_xyz, f->W, f->H, goal_x, goal_z, nullptr);
        FR_CHECK(hipGetLastError());
        FR_CHECK(hipDeviceSynchronize());
        f->gen += 1;
        return true;
    })
}

bool fastrrt_frame_process_safe_zone(void *frame, int min_dist_x, int min_dist_z)
{
    FR_TRY(false, {
        FrFrameObj *f = static_cast<FrFrameObj *>(frame);
        if (!f)
            throw std::runtime_error("null frame");
        if (min_dist_x < 0 || min_dist_z < 0)
            throw std::runtime_error("min_dist must be >= 0");
        uint8_t *near = nullptr;
        FR_CHECK(hipDeviceSynchronize());
        FR_CHECK(hipMalloc(&near, (size_t)f->W * f->H));
        fr_launch_frame_safe_zone(f->d_xyz, f->d_costs, f->n_classes, f->W, f->H, min_dist_x / 2, min_dist_z / 2, near,
                                  nullptr);
        const hipError_t e = hipGetLastError();
        const hipError_t e2 = hipDeviceSynchronize();
        (void)hipFree(near);
        FR_CHECK(e);
        FR_CHECK(e2);
        f->gen += 1;
        return true;
    })
}

/* ---- FastRRT handle ---- */

void *fastrrt_initialize(int width, int height, float perception_width_m, float perception_height_m,
                         float max_steering_angle_deg, float vehicle_length, int lower_bound_x, int lower_bound_z,
                         int upper_bound_x, int upper_bound_z, float *class_costs, float max_curvature)
{
    (void)class_costs; /* unused by the reference too (python_bind.cpp:16-27) */
    FR_TRY(nullptr, {
        return (void *)new FrFastRRT(width, height, perception_width_m, perception_height_m, max_steering_angle_deg,
                                     vehicle_length, lower_bound_x, lower_bound_z, upper_bound_x, upper_bound_z,
                                     max_curvature);
    })
}

void fastrrt_destroy(void *ptr) { FR_TRY_VOID({ delete static_cast<FrFastRRT *>(ptr); }) }

void set_plan_data(void *ptr, void *frame, int start_x, int start_z, float start_heading_rad, int goal_x, int goal_z,
                   float goal_heading_rad, float velocity_m_s, int min_dist_x, int min_dist_z, int timeout_ms,
                   float max_path_size, float dist_to_goal_tolerance, float heading_error_tolerance_rad)
{
    FR_TRY_VOID({
        R(ptr)->set_plan_data(F(frame), start_x, start_z, start_heading_rad, goal_x, goal_z, goal_heading_rad,
                              velocity_m_s, min_dist_x, min_dist_z, timeout_ms, max_path_size, dist_to_goal_tolerance,
                              heading_error_tolerance_rad);
    })
}

bool goal_reached(void *ptr) { FR_TRY(false, { return R(ptr)->goal_reached(); }) }

void search_init(void *ptr, bool copy_intrinsic_costs_from_frame)
{
    (void)copy_intrinsic_costs_from_frame; /* ignored by the reference (fastrrt.cpp:85-98) */
    FR_TRY_VOID({ R(ptr)->search_init(); })
}

bool loop(void *ptr, bool smart_expansion) { FR_TRY(false, { return R(ptr)->loop(smart_expansion); }) }

bool path_optimize(void *ptr) { FR_TRY(false, { return R(ptr)->path_optimize(); }) }

int *export_graph_nodes(void *ptr)
{
    FR_TRY(nullptr, {
        auto nodes = R(ptr)->g.list_all();
        int *res = new int[3 * nodes.size() + 1];
        res[0] = (int)nodes.size();
        for (size_t i = 0; i < nodes.size(); ++i) {
            res[3 * i + 1] = nodes[i].x;
            res[3 * i + 2] = nodes[i].z;
            res[3 * i + 3] = nodes[i].type;
        }
        return res;
    })
}

void release_export_graph_nodes(float *ptr) { delete[] reinterpret_cast<int *>(ptr); }

float *get_planned_path(void *ptr)
{
    FR_TRY(nullptr, { return convert_path(R(ptr)->get_planned_path()); })
}

float *interpolate_planned_path(void *ptr)
{
    FR_TRY(nullptr, {
        FrFastRRT *r = R(ptr);
        auto p = r->get_planned_path();
        return convert_path(interpolate_path(p, r->g.W(), r->g.H()));
    })
}

void release_planned_path_data(float *ptr) { delete[] ptr; }

float *interpolate_planned_path_p(void *ptr, float *p, int size)
{
    FR_TRY(nullptr, {
        FrFastRRT *r = R(ptr);
        std::vector<Waypoint> pref;
        if (size < 0 || (size > 0 && !p))
            throw std::runtime_error("interpolate_planned_path_p: invalid buffer");
        /* whole (x, z, heading) triples only: a trailing partial triple is
         * never read past the caller's `size` floats */
        for (int i = 0; i + 2 < size; i += 3)
            pref.push_back({(int)p[i], (int)p[i + 1], p[i + 2]});
        return convert_path(interpolate_path(pref, r->g.W(), r->g.H()));
    })
}

float *ideal_curve(void *ptr, int goal_x, int goal_z, float goal_heading_rad)
{
    FR_TRY(nullptr, { return convert_path(R(ptr)->ideal_curve(goal_x, goal_z, goal_heading_rad)); })
}

void compute_region_debug_performance(void *ptr) { FR_TRY_VOID({ R(ptr)->g.compute_region_density(); }) }

void save_current_graph_state(void *ptr, const char *filename)
{
    FR_TRY_VOID({ R(ptr)->g.dump(filename); })
}

void load_graph_state(void *ptr, const char *filename)
{
    FR_TRY_VOID({
        FrFastRRT *r = R(ptr);
        r->g.read_dump(filename);
        r->last_count = r->g.count(FR_TYPE_TEMP);
        FrCtl c;
        r->g.eng.read_ctl(0, &c);
        c.last_temp = r->last_count;
        r->g.eng.write_ctl(0, c);
    })
}

void fastrrt_set_seed(void *ptr, uint64_t seed, uint64_t stream)
{
    FR_TRY_VOID({
        R(ptr)->seed = seed;
        R(ptr)->stream = stream;
    })
}

void fastrrt_set_density_region(void *ptr, int side)
{
    FR_TRY_VOID({
        FrFastRRT *r = R(ptr);
        if (side < 16)
            throw std::runtime_error("density region side must be at least 16 px");
        r->g.prm.region_side = side;
        r->g.rebind();
    })
}

int fastrrt_iterations(void *ptr) { FR_TRY(-1, { return R(ptr)->iterations; }) }

void fastrrt_iteration_stats(void *ptr, int *n_expansions, int *n_new, int *n_checked)
{
    FR_TRY_VOID({
        FrFastRRT *r = R(ptr);
        *n_expansions = r->last_expansions;
        *n_new = r->last_new;
        *n_checked = r->last_fresh;
    })
}

/* ---- graph handle ---- */

void *cudagraph_initialize(int width, int height, float perception_width_m, float perception_height_m,
                           float max_steering_angle_deg, float vehicle_length, int min_distance_x, int min_distance_z,
                           int lower_bound_x, int lower_bound_z, int upper_bound_x, int upper_bound_z,
                           float *segmentation_class_cost, float max_curvature)
{
    FR_TRY(nullptr, {
        std::unique_ptr<FrGraph> g(new FrGraph(width, height));
        g->set_search(min_distance_x, min_distance_z, lower_bound_x, lower_bound_z, upper_bound_x, upper_bound_z);
        g->set_physical(perception_width_m, perception_height_m, (double)max_steering_angle_deg * FR_TO_RAD_D,
                        vehicle_length, max_curvature);
        int count = segmentation_class_cost ? (int)segmentation_class_cost[0] : 0;
        if (count < 0)
            count = 0;
        g->set_class_costs(segmentation_class_cost ? segmentation_class_cost + 1 : nullptr, count);
        return (void *)g.release();
    })
}

void cudagraph_destroy(void *ptr) { FR_TRY_VOID({ delete static_cast<FrGraph *>(ptr); }) }

void compute_apf_repulsion(void *ptr, void *frame, float kr, int radius)
{
    FR_TRY_VOID({ G(ptr)->apf_repulsion(F(frame), kr, radius); })
}

bool compute_apf_repulsion_noalloc(void *ptr, void *frame, float kr, int radius)
{
    FR_TRY(false, {
        G(ptr)->apf_repulsion_noalloc(F(frame), kr, radius);
        return true;
    })
}

size_t fastrrt_apf_noalloc_scratch_bytes(int width, int height, int radius)
{
    FR_TRY((size_t)0, { return fr_apf_scratch_bytes(width, height, radius); })
}

size_t fastrrt_apf_noalloc_scratch_capacity(void) { return (size_t)FR_APF_SCRATCH_BYTES; }

void compute_apf_attraction(void *ptr, void *frame, float ka, int goal_x, int goal_z)
{
    FR_TRY_VOID({ G(ptr)->apf_attraction(F(frame), ka, goal_x, goal_z); })
}

float *get_intrinsic_costs(void *ptr)
{
    FR_TRY(nullptr, {
        auto v = G(ptr)->intrinsic_costs();
        float *d = new float[v.size()];
        std::copy(v.begin(), v.end(), d);
        return d;
    })
}

void destroy_intrinsic_costs_ptr(float *ptr) { delete[] ptr; }

void add(void *ptr, int x, int z, float heading_rad, int parent_x, int parent_z, float cost)
{
    FR_TRY_VOID({ G(ptr)->add(x, z, heading_rad, parent_x, parent_z, cost); })
}

void add_temporary(void *ptr, int x, int z, float heading_rad, int parent_x, int parent_z, float cost)
{
    FR_TRY_VOID({ G(ptr)->add_temporary(x, z, heading_rad, parent_x, parent_z, cost); })
}

void derivate_node(void *ptr, void *frame, float steering_angle, int path_size, float velocity, int x, int z)
{
    FR_TRY_VOID({ G(ptr)->derivate(F(frame), steering_angle, (float)(double)path_size, velocity, x, z); })
}

void accept_derived_nodes(void *ptr, int goal_x, int goal_z, float goal_heading)
{
    (void)goal_x;
    (void)goal_z;
    (void)goal_heading;
    FR_TRY_VOID({ G(ptr)->accept_derived(); })
}

bool check_in_graph(void *ptr, int x, int z) { FR_TRY(false, { return G(ptr)->check_in_graph(x, z); }) }

float get_heading(void *ptr, int x, int z) { FR_TRY(0.0f, { return G(ptr)->get_heading(x, z); }) }

void clear(void *ptr) { FR_TRY_VOID({ G(ptr)->clear(); }) }

float *list_all(void *ptr)
{
    FR_TRY(nullptr, {
        FrGraph *g = G(ptr);
        auto nodes = g->list_all();
        std::vector<float> hd(g->N());
        FR_CHECK(hipMemcpyAsync(hd.data(), g->slab.heading, 4 * (size_t)g->N(), hipMemcpyDeviceToHost, g->st()));
        g->eng.sync();
        float *res = new float[nodes.size() * 4 + 1];
        res[0] = 0.0f + (float)nodes.size();
        for (size_t i = 0; i < nodes.size(); ++i) {
            res[4 * i + 1] = (float)nodes[i].x;
            res[4 * i + 2] = (float)nodes[i].z;
            res[4 * i + 3] = hd[nodes[i].z * g->W() + nodes[i].x];
            res[4 * i + 4] = (float)nodes[i].type;
        }
        return res;
    })
}

void free_list_all(float *p) { delete[] p; }

/* The reference converts the curve's float to bool: a failed curve (-1, -2)
 * answers true, only a zero cost or a blocked goal answer false. */
bool can_connect_to_goal_using_hermite(void *ptr, void *frame, int x, int z, int goal_x, int goal_z,
                                       float goal_heading)
{
    FR_TRY(false, {
        const auto r = G(ptr)->can_connect_hermite(F(frame), x, z, goal_x, goal_z, host_trunc((double)goal_heading));
        return !r.first && r.second != 0.0f;
    })
}

int cudagraph_can_connect_hermite(void *ptr, void *frame, int x, int z, int goal_x, int goal_z, float goal_heading,
                                  float *result)
{
    FR_TRY(-1, {
        const auto r = G(ptr)->can_connect_hermite(F(frame), x, z, goal_x, goal_z, host_trunc((double)goal_heading));
        if (result)
            *result = r.second;
        return r.first ? 1 : 0;
    })
}

bool cudagraph_check_feasible_connection(void *ptr, void *frame, int start_x, int start_z, int end_x, int end_z,
                                         int velocity_m_s)
{
    FR_TRY(false, {
        int32_t q[4];
        q[0] = start_x;
        q[1] = start_z;
        q[2] = end_x;
        q[3] = end_z;
        uint8_t ok = 0;
        G(ptr)->feasible_connections(F(frame), q, 1, velocity_m_s, &ok, nullptr, nullptr);
        return ok != 0;
    })
}

int cudagraph_check_feasible_connections(void *ptr, void *frame, const int32_t *queries, int n, int velocity_m_s,
                                         uint8_t *feasible, double *final_heading, float *path_cost)
{
    FR_TRY(-1, {
        G(ptr)->feasible_connections(F(frame), queries, n, velocity_m_s, feasible, final_heading, path_cost);
        return n;
    })
}

void solve_collisions(void *ptr) { FR_TRY_VOID({ G(ptr)->solve_collisions(); }) }

void process_direct_goal_connection(void *ptr, void *frame, int goal_x, int goal_z, float goal_heading_rad,
                                    float max_curvature)
{
    FR_TRY_VOID({ G(ptr)->process_direct_goal_connection(F(frame), goal_x, goal_z, goal_heading_rad, max_curvature); })
}

bool is_directly_connected_to_goal(void *ptr, int x, int z) { FR_TRY(false, { return G(ptr)->gc_flag(x, z); }) }

float direct_connection_to_goal_cost(void *ptr, int x, int z)
{
    FR_TRY(-1.0f, {
        FrGraph *g = G(ptr);
        if (!g->gc_flag(x, z))
            return -1.0f;
        return g->gc_cost(x, z);
    })
}

float direct_connection_to_goal_heading(void *ptr, int x, int z)
{
    FR_TRY(0.0f, { return G(ptr)->gc_heading(x, z); })
}

void dump_graph_to_file(void *ptr, char *filename) { FR_TRY_VOID({ G(ptr)->dump(filename); }) }

void read_from_dump_file(void *ptr, char *filename) { FR_TRY_VOID({ G(ptr)->read_dump(filename); }) }

int *get_parent(void *ptr, int x, int z)
{
    FR_TRY(nullptr, {
        auto p = G(ptr)->get_parent(x, z);
        int *d = new int[2];
        d[0] = p.first;
        d[1] = p.second;
        return d;
    })
}

void free_parent_data(int *data_ptr) { delete[] data_ptr; }

float get_cost(void *ptr, int x, int z) { FR_TRY(-1.0f, { return G(ptr)->get_cost(x, z); }) }

int count_all(void *ptr) { FR_TRY(-1, { return G(ptr)->count_all(); }) }

void cudagraph_set_collision(void *ptr, int x, int z, int new_parent_x, int new_parent_z, float heading, float cost)
{
    FR_TRY_VOID({ G(ptr)->set_collision(x, z, new_parent_x, new_parent_z, heading, cost); })
}

int cudagraph_get_type(void *ptr, int x, int z) { FR_TRY(-1, { return G(ptr)->get_type(x, z); }) }

int cudagraph_get_child_count(void *ptr, int x, int z) { FR_TRY(0, { return G(ptr)->child_count(x, z); }) }

int cudagraph_count(void *ptr, int type) { FR_TRY(-1, { return G(ptr)->count(type); }) }

bool cudagraph_check_consistent(void *ptr) { FR_TRY(false, { return G(ptr)->check_consistent(); }) }

void cudagraph_find_best_node(void *ptr, void *frame, float heading, float radius, int x, int z,
                              float max_heading_error, int *out)
{
    FR_TRY_VOID({
        auto p = G(ptr)->find_best_node(frame ? F(frame) : nullptr, heading, radius, x, z, max_heading_error, false);
        out[0] = p.first;
        out[1] = p.second;
    })
}

bool cudagraph_check_goal_reached(void *ptr, int goal_x, int goal_z, float heading, float tolerance,
                                  float max_heading_error)
{
    FR_TRY(false, { return G(ptr)->check_goal_reached(goal_x, goal_z, heading, tolerance, max_heading_error); })
}

void cudagraph_derivate_node(void *ptr, void *frame, float steering_angle, float path_size, float velocity, int x,
                             int z, int *out)
{
    FR_TRY_VOID({
        auto p = G(ptr)->derivate(F(frame), steering_angle, path_size, velocity, x, z);
        out[0] = p.first;
        out[1] = p.second;
    })
}

void fastrrt_export_state(void *ptr, int32_t *parent_x, int32_t *parent_z, int32_t *type, int32_t *child_count,
                          float *heading, float *cost, uint8_t *goal_flag, float *goal_heading, float *goal_cost)
{
    FR_TRY_VOID({
        R(ptr)->g.export_state(parent_x, parent_z, type, child_count, heading, cost, goal_flag, goal_heading,
                               goal_cost);
    })
}

void cudagraph_export_state(void *ptr, int32_t *parent_x, int32_t *parent_z, int32_t *type, int32_t *child_count,
                            float *heading, float *cost, uint8_t *goal_flag, float *goal_heading, float *goal_cost)
{
    FR_TRY_VOID({
        G(ptr)->export_state(parent_x, parent_z, type, child_count, heading, cost, goal_flag, goal_heading, goal_cost);
    })
}

void fastrrt_profile_enable(bool on) { FR_TRY_VOID({ fr_prof_enable(on); }) }
void fastrrt_profile_split(bool on) { FR_TRY_VOID({ fr_prof_split(on); }) }

void fastrrt_profile_reset(void) { FR_TRY_VOID({ fr_prof_reset(); }) }

bool fastrrt_profile_read(int kernel, long *launches, double *total_ms)
{
    FR_TRY(false, {
        if (kernel < 0 || kernel >= FR_PROF_N)
            throw std::runtime_error("unknown profiled kernel");
        fr_prof_read(kernel, launches, total_ms);
        return true;
    })
}

} /* extern "C" */
