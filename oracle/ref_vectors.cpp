// ref_vectors.cpp -- driver appended to the reference's host-only code by
// oracle/ref_build.sh (same translation unit).  It calls the reference's own
// functions and prints JSON golden vectors on stdout:
//   ref_vectors plan                   seeded_sector_value, rowmix_realized_seed,
//                                      generate_rowmix_plan, rowmix_plan_hash
//   ref_vectors state                  the mt19937_64 draws of run_eliminated_sector
//   ref_vectors apply INSTANCE SECTOR  load_cyclic_instance + sector layouts +
//                                      cpu_eliminated_forward (F Dc x) and
//                                      cpu_apply_target_preconditioner (P^T Dr P)
//                                      on the reference's deterministic blocks
// Test infrastructure only: nothing in the product links or runs it.

static void put_u64_list(const std::vector<uint64_t>& v) {
    std::cout << "[";
    for (size_t i = 0; i < v.size(); ++i) std::cout << (i ? ", " : "") << v[i];
    std::cout << "]";
}

static uint64_t fnv_u32_values(const std::vector<uint64_t>& v) {
    uint64_t h = FNV1A64_OFFSET;
    h = fnv1a64_append_u64(h, v.size());
    for (uint64_t x : v) h = fnv1a64_append_u64(h, x);
    return h;
}

static int cmd_plan() {
    // realized seeds, plan hashes and pair samples at the genus-12 and genus-20
    // sector heights for the paper primes and 65521
    const std::vector<std::string> ops = {"eliminated", "augmented", "replacement"};
    const std::vector<size_t> heights = {104, 106, 21872, 21886};
    const std::vector<uint32_t> primes = {109, 661, 65521};
    std::cout << "{\"seeded_sector_value\": [";
    bool first = true;
    for (uint64_t seed : {0ULL, 1ULL, 10001ULL, 0xFFFFFFFFFFFFFFFFULL})
        for (size_t sector : {0, 1, 5, 11}) {
            std::cout << (first ? "" : ", ") << "[" << seed << ", " << sector << ", "
                      << seeded_sector_value(seed, sector) << "]";
            first = false;
        }
    std::cout << "], \"plans\": [";
    first = true;
    for (const std::string& op : ops)
        for (size_t rows : heights)
            for (uint32_t p : primes)
                for (size_t sector : {0, 3}) {
                    RowMixOptions opt{2, 10001};
                    RowMixPlan plan = generate_rowmix_plan(rows, op, sector, p, opt);
                    std::vector<uint64_t> head, tail;
                    for (size_t r = 0; r < plan.round_pairs.size(); ++r) {
                        const auto& pr = plan.round_pairs[r];
                        for (size_t q = 0; q < 3 && q < pr.size(); ++q)
                            head.insert(head.end(), {pr[q].first, pr[q].second, pr[q].a, pr[q].b});
                        if (!pr.empty()) {
                            const auto& e = pr.back();
                            tail.insert(tail.end(), {e.first, e.second, e.a, e.b});
                        }
                    }
                    std::cout << (first ? "" : ", ") << "{\"op\": \"" << op << "\", \"rows\": " << rows
                              << ", \"prime\": " << p << ", \"sector\": " << sector
                              << ", \"rounds\": 2, \"seed\": 10001, \"realized_seed\": "
                              << plan.realized_seed << ", \"plan_hash\": " << rowmix_plan_hash(plan)
                              << ", \"first_pairs\": ";
                    put_u64_list(head);
                    std::cout << ", \"last_pairs\": ";
                    put_u64_list(tail);
                    std::cout << "}";
                    first = false;
                }
    std::cout << "]}" << std::endl;
    return 0;
}

static int cmd_state() {
    // Dr, Dc, V of run_eliminated_sector (cuprym_cyclic.cu:2196-2199) for small shapes
    std::cout << "{\"states\": [";
    bool first = true;
    for (uint32_t p : {109u, 661u, 65521u})
        for (size_t sector : {0, 4}) {
            const size_t rows = 37, cols = 29, b = 4;
            std::mt19937_64 gen(seeded_sector_value(1, sector));
            std::vector<uint16_t> dr = generate_random_vector<uint16_t>(rows, (uint16_t)p, gen, true);
            std::vector<uint16_t> dc = generate_random_vector<uint16_t>(cols, (uint16_t)p, gen, true);
            std::vector<uint16_t> v = generate_random_vector<uint16_t>(cols * b, (uint16_t)p, gen);
            std::cout << (first ? "" : ", ") << "{\"prime\": " << p << ", \"seed\": 1, \"sector\": "
                      << sector << ", \"rows\": " << rows << ", \"cols\": " << cols << ", \"b\": " << b;
            std::cout << ", \"dr\": ";
            put_u64_list(std::vector<uint64_t>(dr.begin(), dr.end()));
            std::cout << ", \"dc\": ";
            put_u64_list(std::vector<uint64_t>(dc.begin(), dc.end()));
            std::cout << ", \"v\": ";
            put_u64_list(std::vector<uint64_t>(v.begin(), v.end()));
            std::cout << "}";
            first = false;
        }
    std::cout << "]}" << std::endl;
    return 0;
}

// The set-up of run_eliminated_sector (cuprym_cyclic.cu:2136-2257, main
// :2617-2680), then the forward half of cpu_apply_eliminated_square
// (:1921-1950) on the three salted blocks of validate_eliminated_cpu_cuda
// (:2064-2083) with 4 columns: F (Dc x) and P^T Dr P F (Dc x).
static int cmd_apply(const std::string& path, size_t sector) {
    CyclicInstance instance = load_cyclic_instance(path);
    CharacterMap a0_map = build_character_map(instance.weight_a0, instance.r, "A0");
    CharacterMap a1_map = build_character_map(instance.weight_a1, instance.r, "A1");
    std::vector<size_t> u_indices;
    std::vector<uint16_t> u_weights;
    for (size_t i = 0; i < instance.n; ++i) {
        if (i == instance.elimination_w_index) continue;
        u_indices.push_back(i);
        u_weights.push_back(instance.weight_v[i]);
    }
    CharacterMap kernel_map = build_character_map(instance.elimination_kernel_weights, instance.r, "ker(mu_w)");
    CyclicCommonIncidence common_dm = build_common_incidences_generic(
        u_indices.size(), instance.m, instance.r, u_indices, u_weights, "eliminated D_m");
    CyclicCommonIncidence common_d2 = build_common_incidences_generic(
        u_indices.size(), instance.m - 1, instance.r, u_indices, u_weights, "eliminated D_(m-1)");
    size_t y_sector = sub_mod(sector, instance.elimination_w_weight, instance.r);
    CyclicSectorLayout dm_layout = build_sector_layout_generic(instance, common_dm, a0_map, a1_map, sector, "eliminated D_m");
    CyclicSectorLayout d2_layout = build_sector_layout_generic(instance, common_d2, a0_map, a1_map, y_sector, "eliminated D_(m-1)");
    CyclicTensorLayout z_layout = build_tensor_layout(common_dm.row_subset_weights, kernel_map, instance.r, y_sector, "eliminated z");
    size_t x_cols = dm_layout.column_descriptors.size();
    size_t t_rows = dm_layout.row_descriptors.size();
    size_t y_rows = d2_layout.column_descriptors.size();
    size_t z_cols = z_layout.descriptors.size();
    size_t target_rows = d2_layout.row_descriptors.size();
    size_t total_cols = x_cols + z_cols;
    my_int storage_prime = static_cast<my_int>(instance.modulus);
    std::mt19937_64 gen(seeded_sector_value(1, sector));
    vector<my_int> row_precond = generate_random_vector(target_rows, storage_prime, gen, true);
    vector<my_int> col_precond = generate_random_vector(total_cols, storage_prime, gen, true);
    RowMixOptions opt{2, 10001};
    RowMixPlan rowmix_plan = generate_rowmix_plan(target_rows, "eliminated", sector, instance.modulus, opt);
    vector<my_int> d2_row_precond_for_operator = rowmix_plan.enabled() ? ones_vector(target_rows) : row_precond;
    CyclicPhiHostData<my_int> dm_host = make_operator_host_data(
        instance, common_dm, std::move(dm_layout), a0_map, a1_map, ones_vector(t_rows), ones_vector(x_cols));
    CyclicPhiHostData<my_int> d2_host = make_operator_host_data(
        instance, common_d2, std::move(d2_layout), a0_map, a1_map, std::move(d2_row_precond_for_operator),
        ones_vector(y_rows));
    const size_t dense_cols = 4;
    std::cout << "{\"instance\": \"" << path << "\", \"sector\": " << sector << ", \"rows\": " << target_rows
              << ", \"cols\": " << total_cols << ", \"prime\": " << instance.modulus << ", \"b\": " << dense_cols
              << ", \"plan_hash\": " << rowmix_plan_hash(rowmix_plan) << ", \"applies\": [";
    bool first = true;
    for (uint64_t salt : {0x404ULL, 0x505ULL, 0x606ULL}) {
        vector<my_int> input = deterministic_dense_block(total_cols, dense_cols, instance.modulus, salt);
        vector<my_int> target;
        cpu_eliminated_forward(dm_host, d2_host, z_layout, instance, col_precond, x_cols, z_cols, input,
                               dense_cols, instance.modulus, target);
        std::vector<uint64_t> in64(input.begin(), input.end()), f64(target.begin(), target.end());
        cpu_apply_target_preconditioner(rowmix_plan, row_precond, target_rows, dense_cols, instance.modulus,
                                        target);
        std::vector<uint64_t> m64(target.begin(), target.end());
        std::cout << (first ? "" : ", ") << "{\"salt\": " << salt << ", \"input_fnv1a64\": " << fnv_u32_values(in64)
                  << ", \"forward_fnv1a64\": " << fnv_u32_values(f64) << ", \"mixed_fnv1a64\": "
                  << fnv_u32_values(m64) << ", \"forward_head\": ";
        put_u64_list(std::vector<uint64_t>(f64.begin(), f64.begin() + std::min<size_t>(16, f64.size())));
        std::cout << ", \"mixed_head\": ";
        put_u64_list(std::vector<uint64_t>(m64.begin(), m64.begin() + std::min<size_t>(16, m64.size())));
        std::cout << "}";
        first = false;
    }
    std::cout << "]}" << std::endl;
    return 0;
}

int main(int argc, char** argv) {
    try {
        const std::string cmd = argc > 1 ? argv[1] : "";
        if (cmd == "plan") return cmd_plan();
        if (cmd == "state") return cmd_state();
        if (cmd == "apply" && argc == 4) return cmd_apply(argv[2], std::stoul(argv[3]));
        std::cerr << "usage: ref_vectors plan | state | apply INSTANCE SECTOR" << std::endl;
        return 1;
    } catch (const std::exception& e) {
        std::cerr << "ref_vectors error: " << e.what() << std::endl;
        return 1;
    }
}
