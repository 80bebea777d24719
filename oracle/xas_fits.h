// GENERATED by oracle/gen_fits.py from the reference's
// src/waveformx/models/xpnr/_generated/{utilities,qnm,inspiral,intermediate,ringdown}.py
// (LALSimulation 6.2.0's IMRPhenomX fits via waveformrs; GPL-3.0-or-later data).
// Oracle only: generic over the scalar R (double, or the oracle's forward-mode dual),
// fma(a, b, c) unfused as a * b + c. Do not edit.
#pragma once

// ---- utilities.py ----
// utilities.py:29
template <class R> static inline R ut_sign(R x) {
    return o_where((x > 0.0), R(1.0), o_where((x < 0.0), R((-1.0)), R(0.0)));
}

// utilities.py:34
template <class R> static inline R ut_chi_eff(R eta, R chi1l, R chi2l) {
    R delta = o_sqrt_nonneg(o_fma((-4.0), eta, 1.0));
    R mm1 = (0.5 * (1.0 + delta));
    R mm2 = (0.5 * (1.0 - delta));
    return o_fma(mm1, chi1l, (mm2 * chi2l));
}

// utilities.py:42
template <class R> static inline R ut_chi_pn_hat(R eta, R chi1l, R chi2l) {
    R delta = o_sqrt_nonneg(o_fma((-4.0), eta, 1.0));
    R mm1 = (0.5 * (1.0 + delta));
    R mm2 = (0.5 * (1.0 - delta));
    R chi_eff = o_fma(mm1, chi1l, (mm2 * chi2l));
    return (o_fma((-((38.0 / 113.0) * eta)), (chi1l + chi2l), chi_eff) / (1.0 - ((76.0 * eta) / 113.0)));
}

// utilities.py:51
template <class R> static inline R ut_s_tot_r(R eta, R chi1l, R chi2l) {
    R delta = o_sqrt_nonneg(o_fma((-4.0), eta, 1.0));
    R m1 = (0.5 * (1.0 + delta));
    R m2 = (0.5 * (1.0 - delta));
    R m1s = (m1 * m1);
    R m2s = (m2 * m2);
    return (o_fma(m1s, chi1l, (m2s * chi2l)) / (m1s + m2s));
}

// utilities.py:61
template <class R> static inline R ut_dchi(R chi1l, R chi2l) {
    return (chi1l - chi2l);
}

// utilities.py:16
template <class R> static inline R ut_f_isco(R chif) {
    R z1 = o_fma(o_cbrt(o_fma((-chif), chif, 1.0)), (o_cbrt((1.0 + chif)) + o_cbrt((1.0 - chif))), 1.0);
    z1 = o_min(z1, R(3.0));
    R z2 = o_sqrt(o_fma((3.0 * chif), chif, (z1 * z1)));
    R r_isco = o_fma((-ut_sign(chif)), o_sqrt(((3.0 - z1) * o_fma(2.0, z2, (3.0 + z1)))), (3.0 + z2));
    R r_isco_sq = o_sqrt(r_isco);
    R r_isco_3o2 = ((r_isco_sq * r_isco_sq) * r_isco_sq);
    R omega_isco = (1.0 / (r_isco_3o2 + chif));
    return (omega_isco / M_PI_);
}

// utilities.py:66
template <class R> static inline R ut_f_meco(R eta, R chi1l, R chi2l) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R eta4 = (eta3 * eta);
    R delta = o_sqrt_nonneg(o_fma((-4.0), eta, 1.0));
    R s = ut_chi_pn_hat(eta, chi1l, chi2l);
    R s2 = (s * s);
    R s3 = (s2 * s);
    R dchi = (chi1l - chi2l);
    R dchi2 = (dchi * dchi);
    R no_spin = (o_fma((-6.693930988501673e-05), eta3, o_fma(0.003940354686136861, eta2, o_fma(0.0077903147004616865, eta, 0.018744340279608845))) / o_fma((-0.10423384680638834), eta, 1.0));
    R eq_spin = ((s * o_fma((-7.547517256664526e-05), s2, o_fma(eta3, o_fma(4.038390455349854e-06, s2, o_fma(3.0679548774047616e-05, s, (-0.0005418851224505745))), o_fma(eta, o_fma((-0.00035182990586857726), s2, o_fma((-1.049013062611254e-05), s, 0.00024422562796266645)), o_fma(eta2, o_fma((-0.00038216081981505285), s2, o_fma((-1.0021140796150737e-05), s, 0.00036556167661117023)), o_fma(eta4, o_fma((-0.0230825153005985), s2, o_fma(0.022386313074011715, s, (-0.0006807631931297156))), o_fma((-2.585252361022052e-05), s, 0.00027180386951683135))))))) / o_fma(o_fma((-0.00767768214056772), eta3, o_fma(0.012622225233586283, eta2, o_fma(1.0, eta4, o_fma(0.014176504653145768, eta, 0.003507640638496499)))), s3, o_fma(o_fma((-0.001956646166089053), eta3, o_fma((-0.019641378027236502), eta2, o_fma((-2.551566872093786), eta4, o_fma((-0.02876774751921441), eta, (-0.0058684526275074025))))), s2, o_fma(o_fma(0.005146809717492324, eta3, o_fma(0.025066696514373803, eta2, o_fma(1.4861197211952053, eta4, o_fma((-0.012429476486138982), eta, (-0.014590539285641243))))), s, 0.026666543809890402))));
    R uneq_spin = o_fma((dchi2 * o_fma(1.6343732281057392e-05, eta, 0.00034375176678815234)), eta2, (((dchi * delta) * eta) * o_fma((-0.0011713735642446144), s, o_fma(0.08064665214195679, eta2, (eta * o_fma((-0.005746537021035632), s, (-0.028476219509487793)))))));
    return ((no_spin + eq_spin) + uneq_spin);
}

// utilities.py:101
template <class R> static inline R ut_final_mass_2017(R eta, R chi1l, R chi2l) {
    R delta = o_sqrt_nonneg(o_fma((-4.0), eta, 1.0));
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R eta4 = (eta3 * eta);
    R s = ut_s_tot_r(eta, chi1l, chi2l);
    R s2 = (s * s);
    R s3 = (s2 * s);
    R dchi = (chi1l - chi2l);
    R dchi2 = (dchi * dchi);
    R no_spin = o_fma(3.145145224278187, eta4, o_fma((-0.84667563764404), eta3, o_fma(0.057190958417936644, eta, (0.5609904135313374 * eta2))));
    R eq_spin = ((o_fma(3.145145224278187, eta4, o_fma((-0.84667563764404), eta3, o_fma(0.057190958417936644, eta, (0.5609904135313374 * eta2)))) * o_fma(o_fma((-10.023747993978121), eta2, o_fma(4.952698546796005, eta, (-0.6320191645391563))), s3, o_fma(o_fma(2.176667900182948, eta2, (-0.17762802148331427)), s2, o_fma(o_fma(5.49074464410971, eta2, o_fma((-1.1387311580238488), eta, (-0.13084389181783257))), s, 1.0)))) / o_fma(o_fma(4.274567337924067, eta2, o_fma(0.367620218664352, eta, (-0.9919475346968611))), s, 1.0));
    eq_spin = (eq_spin - no_spin);
    R uneq_spin = o_fma((-((((0.01978238971523653 * dchi) * delta) * o_fma((-4.91667749015812), eta, 1.0)) * eta)), s, o_fma(((((-0.09803730445895877) * dchi) * delta) * o_fma((-3.2283713377939134), eta, 1.0)), eta2, ((0.01118530335431078 * dchi2) * eta3)));
    return (1.0 - ((no_spin + eq_spin) + uneq_spin));
}

// utilities.py:119
template <class R> static inline R ut_final_spin_2017(R eta, R chi1l, R chi2l) {
    R delta = o_sqrt_nonneg(o_fma((-4.0), eta, 1.0));
    R m1 = (0.5 * (1.0 + delta));
    R m2 = (0.5 * (1.0 - delta));
    R m1sq = (m1 * m1);
    R m2sq = (m2 * m2);
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R s = ut_s_tot_r(eta, chi1l, chi2l);
    R s2 = (s * s);
    R s3 = (s2 * s);
    R dchi = (chi1l - chi2l);
    R dchi2 = (dchi * dchi);
    R no_spin = (o_fma((-12.333573402277912), eta3, o_fma(3.4641016151377544, eta, (20.0830030082033 * eta2))) / o_fma(7.2388440419467335, eta, 1.0));
    R eq_spin = o_fma((m1sq + m2sq), s, (o_fma(o_fma(1.9595643107593743, eta3, o_fma(0.142443244743048, eta, (-(0.9598353840147513 * eta2)))), s3, o_fma(o_fma(1.668810429851045, eta3, o_fma((-0.8561951310209386), eta, (-(0.09939065676370885 * eta2)))), s, (o_fma(3.4768263932898678, eta3, o_fma(0.5881660363307388, eta, (-(2.149269067519131 * eta2)))) * s2))) / o_fma(o_fma((-9.710576749140989), eta3, o_fma(2.3191363426522633, eta, (-0.9142232693081653))), s, 1.0)));
    R uneq_spin = o_fma(((((2.3170397514509933 * dchi) * delta) * o_fma((-3.2624649875884852), eta, 1.0)) * eta3), s, o_fma((((0.3223660562764661 * dchi) * delta) * o_fma(9.332575956437443, eta, 1.0)), eta2, (-((0.059808322561702126 * dchi2) * eta3))));
    return ((no_spin + eq_spin) + uneq_spin);
}

// utilities.py:139
template <class R> static inline R ut_linb(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R eta4 = (eta3 * eta);
    R eta5 = (eta3 * eta2);
    R eta6 = (eta4 * eta2);
    R s2 = (s * s);
    R s3 = (s2 * s);
    R s4 = (s3 * s);
    R no_spin = o_fma((-5789128.656876938), eta6, o_fma(5303591.1346921865, eta5, o_fma((-1922385.1649491738), eta4, o_fma(347213.65466875216, eta3, o_fma((-32243.28428870599), eta2, o_fma(1257.9949740608242, eta, 3155.1635543201924))))));
    R eq_spin = o_fma(o_fma(1320.102640190539, eta2, o_fma((-325.92980012408367), eta, (-12.814828278938885))), s4, o_fma(o_fma(1941.8836639529036, eta2, o_fma((-646.3410199799737), eta, 23.404604124552)), s3, o_fma(o_fma((-380.19778216022763), eta2, o_fma(115.49264174560281, eta, (-24.181508118588667))), s, (o_fma(725.6024119989094, eta2, o_fma((-328.3762360751952), eta, 24.72585609641552)) * s2))));
    R uneq_spin = ((((-148.17317525117338) * dchi) * delta) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// utilities.py:155
template <class R> static inline R ut_psi4_to_strain(R eta, R s, R dchi) {
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = o_fma(16026.897939722587, eta4, o_fma((-9862.84178637907), eta3, o_fma(2097.425116152503, eta2, o_fma((-175.42481512989315), eta, 13.39320482758057))));
    R eq_spin = o_fma(o_fma((-764.5242164872267), eta2, o_fma(149.30405703643288, eta, 13.097315867845788)), s4, o_fma(o_fma((-755.9566456906406), eta2, o_fma(137.33317057388916, eta, 15.649521097877374)), s3, o_fma(o_fma(609.5575850476959, eta2, o_fma((-163.04871764530466), eta, 4.7895602776763)), s, (o_fma(376.9200932531847, eta2, o_fma((-97.51812681228478), eta, 1.3934428041390161)) * s2))));
    R uneq_spin = (((105.37711654943146 * dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// ---- qnm.py ----
// qnm.py:24
template <class R> static inline R qnm_fring22(R af) {
    R x2 = (af * af);
    R x3 = (x2 * af);
    R x4 = (x2 * x2);
    R x5 = (x3 * x2);
    R x6 = (x3 * x3);
    R x7 = (x4 * x3);
    R numerator = o_fma(0.0005821983163192694, x7, o_fma(0.0027489038393367993, x6, o_fma((-0.005852438240997211), x5, o_fma((-0.02501704155363241), x4, o_fma(0.02260924869042963, x3, o_fma(0.09535606290986028, x2, o_fma((-0.14989771215394762), af, 0.05947169566573468)))))));
    R denominator = o_fma(0.0873798215084077, x6, o_fma((-0.6036964688511505), x4, o_fma(2.373335413978394, x2, o_fma((-2.8570126619966296), af, 1.0))));
    return (numerator / denominator);
}

// qnm.py:37
template <class R> static inline R qnm_fdamp22(R af) {
    R x2 = (af * af);
    R x3 = (x2 * af);
    R x4 = (x2 * x2);
    R x5 = (x3 * x2);
    R x6 = (x3 * x3);
    R numerator = o_fma(0.0001504546201236794, x6, o_fma((-0.00014745235759327472), x5, o_fma((-0.004843996907020524), x4, o_fma(0.0008490933750566702, x3, o_fma(0.026822526296575368, x2, o_fma((-0.036989395871554566), af, 0.014158792290965177))))));
    R denominator = o_fma(0.009002719412204133, x6, o_fma((-0.31416610693042507), x4, o_fma(1.8952576220623967, x2, o_fma((-2.5900842798681376), af, 1.0))));
    return (numerator / denominator);
}

// ---- inspiral.py ----
// inspiral.py:16
template <class R> static inline R ins_amp_22_v2(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R eta4 = (eta3 * eta);
    R s2 = (s * s);
    R s3 = (s2 * s);
    R no_spin = (o_fma(0.4845148547154606, eta2, o_fma((-0.06098548699809163), eta, (-0.015178276424448592))) / o_fma(0.09799277215675059, eta, 1.0));
    R eq_spin = (o_fma(o_fma((-0.14404522791467844), eta2, 0.01761591799745109), s2, o_fma(o_fma(0.10495263104245876, eta2, 0.02300153747158323), s, ((o_fma((-0.14189350657140673), eta, 0.04834642258922544) * eta) * s3))) / o_fma((-0.7340448493183307), s, 1.0));
    R uneq_spin = (((dchi * delta) * eta4) * o_fma(34.90874132485147, eta, 0.0018724905795891192));
    return ((no_spin + eq_spin) + uneq_spin);
}

// inspiral.py:29
template <class R> static inline R ins_amp_22_v3(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R eta4 = (eta3 * eta);
    R no_spin = (o_fma(8.4630293045015, eta2, o_fma((-1.1970535595488723), eta, (-0.058572000924124644))) / o_fma(15.430818840453686, eta, 1.0));
    R eq_spin = ((o_fma((-0.018924013869130434), s, o_fma(eta2, o_fma(0.8282888482429105, s, 0.788717372588848), o_fma(eta, o_fma((-0.21291764491897636), s, (-0.20646621646484237)), (-0.08746408292050666)))) * s) / o_fma(1.0, s, (-1.332123330797879)));
    R uneq_spin = (((dchi * delta) * eta4) * o_fma(105.84553997647659, eta, 0.004389995099201855));
    return ((no_spin + eq_spin) + uneq_spin);
}

// inspiral.py:40
template <class R> static inline R ins_amp_22_v4(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R eta4 = (eta3 * eta);
    R s2 = (s * s);
    R no_spin = (o_fma(5.629598195000046, eta3, o_fma((-3.186012733446088), eta2, o_fma(1.617404703616985, eta, (-0.16212854591357853)))) / o_fma(0.04507019231274476, eta, 1.0));
    R eq_spin = ((s * o_fma(eta3, o_fma(19.099315016873643, s, (-41.0378120175805)), o_fma(eta, o_fma(5.215501942120774, s, (-4.127597118865669)), o_fma((-0.31443470118113853), s, o_fma(eta2, o_fma((-18.80590889704093), s, 18.353433894421833), 1.0055835408962206))))) / o_fma(1.0, s2, o_fma((-5.717874483424523), s, 5.852706459485663)));
    R uneq_spin = (((dchi * delta) * eta4) * o_fma(208.92352600701068, eta, 0.05575955418803233));
    return ((no_spin + eq_spin) + uneq_spin);
}

// inspiral.py:52
template <class R> static inline R ins_pow_7_of(R x) {
    R pow2 = (x * x);
    return (((pow2 * pow2) * pow2) * x);
}

// inspiral.py:58
template <class R> static inline R ins_amp_22_rho1(R v1, R v2, R v3, R f1, R f2, R f3) {
    R f1p1o3 = o_cbrt(f1);
    R f2p1o3 = o_cbrt(f2);
    R f3p1o3 = o_cbrt(f3);
    R f1p7o3 = ins_pow_7_of(f1p1o3);
    R f1p8o3 = (f1p7o3 * f1p1o3);
    R f2p7o3 = ins_pow_7_of(f2p1o3);
    R f2p8o3 = (f2p7o3 * f2p1o3);
    R f3p7o3 = ins_pow_7_of(f3p1o3);
    R f3p8o3 = (f3p7o3 * f3p1o3);
    return (o_fma((((f1 * f1) * f1) * f2p8o3), v3, o_fma((-(f1p8o3 * ((f2 * f2) * f2))), v3, o_fma((-(((f1 * f1) * f1) * f3p8o3)), v2, o_fma((f1p8o3 * ((f3 * f3) * f3)), v2, o_fma((-(f2p8o3 * ((f3 * f3) * f3))), v1, ((((f2 * f2) * f2) * f3p8o3) * v1)))))) / (((((f1p7o3 * (f1p1o3 - f2p1o3)) * f2p7o3) * (f1p1o3 - f3p1o3)) * (f2p1o3 - f3p1o3)) * f3p7o3));
}

// inspiral.py:72
template <class R> static inline R ins_amp_22_rho2(R v1, R v2, R v3, R f1, R f2, R f3) {
    R f1p1o3 = o_cbrt(f1);
    R f2p1o3 = o_cbrt(f2);
    R f3p1o3 = o_cbrt(f3);
    R f1p7o3 = ins_pow_7_of(f1p1o3);
    R f2p7o3 = ins_pow_7_of(f2p1o3);
    R f3p7o3 = ins_pow_7_of(f3p1o3);
    return (o_fma((-(((f1 * f1) * f1) * f2p7o3)), v3, o_fma((f1p7o3 * ((f2 * f2) * f2)), v3, o_fma((((f1 * f1) * f1) * f3p7o3), v2, o_fma((-(f1p7o3 * ((f3 * f3) * f3))), v2, o_fma((f2p7o3 * ((f3 * f3) * f3)), v1, (-((((f2 * f2) * f2) * f3p7o3) * v1))))))) / (((((f1p7o3 * (f1p1o3 - f2p1o3)) * f2p7o3) * (f1p1o3 - f3p1o3)) * (f2p1o3 - f3p1o3)) * f3p7o3));
}

// inspiral.py:83
template <class R> static inline R ins_amp_22_rho3(R v1, R v2, R v3, R f1, R f2, R f3) {
    R f1p1o3 = o_cbrt(f1);
    R f2p1o3 = o_cbrt(f2);
    R f3p1o3 = o_cbrt(f3);
    R f1p7o3 = ins_pow_7_of(f1p1o3);
    R f1p8o3 = (f1p7o3 * f1p1o3);
    R f2p7o3 = ins_pow_7_of(f2p1o3);
    R f2p8o3 = (f2p7o3 * f2p1o3);
    R f3p7o3 = ins_pow_7_of(f3p1o3);
    R f3p8o3 = (f3p7o3 * f3p1o3);
    return (o_fma((-(f1p7o3 * f2p8o3)), v3, o_fma((f1p8o3 * f2p7o3), v3, o_fma((f1p7o3 * f3p8o3), v2, o_fma((-(f1p8o3 * f3p7o3)), v2, o_fma((f2p8o3 * f3p7o3), v1, (-((f2p7o3 * f3p8o3) * v1))))))) / (((((f1p7o3 * (f1p1o3 - f2p1o3)) * f2p7o3) * (f1p1o3 - f3p1o3)) * (f2p1o3 - f3p1o3)) * f3p7o3));
}

// inspiral.py:104
template <class R> static inline R ins_amp_22_ansatz(R mf, const Powers<R>& powers, const AmpCoeffs<R>& amp) {
    return (((amp.pn_initial + (powers.two_thirds * amp.pn_two_thirds)) + (mf * amp.pn_three_thirds)) + (mf * ((((powers.one_third * amp.pn_four_thirds) + (powers.two_thirds * amp.pn_five_thirds)) + (mf * amp.pn_six_thirds)) + (mf * (((powers.one_third * amp.rho1) + (powers.two_thirds * amp.rho2)) + (mf * amp.rho3))))));
}

// inspiral.py:109
template <class R> static inline R ins_amp_22_ansatz_exact(R mf, const Powers<R>& powers, const AmpCoeffs<R>& amp) {
    R inner = o_fma(powers.one_third, amp.rho1, (powers.two_thirds * amp.rho2));
    inner = o_fma(mf, amp.rho3, inner);
    R middle = o_fma(powers.one_third, amp.pn_four_thirds, (powers.two_thirds * amp.pn_five_thirds));
    middle = o_fma(mf, amp.pn_six_thirds, middle);
    middle = o_fma(mf, inner, middle);
    R pn_amp = o_fma(powers.one_third, amp.pn_one_third, amp.pn_initial);
    pn_amp = o_fma(powers.two_thirds, amp.pn_two_thirds, pn_amp);
    pn_amp = o_fma(mf, amp.pn_three_thirds, pn_amp);
    return o_fma(mf, middle, pn_amp);
}

// inspiral.py:122
template <class R> static inline R ins_amp_22_dansatz(R mf, const WfVars<R>& wf, const AmpCoeffs<R>& amp) {
    R eta = wf.eta;
    R chi1l = wf.chi1l;
    R chi2l = wf.chi2l;
    R rho1 = amp.rho1;
    R rho2 = amp.rho2;
    R rho3 = amp.rho3;
    R chi1l2 = (chi1l * chi1l);
    R chi1l3 = (chi1l2 * chi1l);
    R chi2l2 = (chi2l * chi2l);
    R eta2 = (eta * eta);
    R eta3 = (eta * eta2);
    R mf2 = (mf * mf);
    R lalpi = M_PI_;
    R delta = wf.delta;
    return o_fma((3.0 * mf2), rho3, o_fma(((8.0 / 3.0) * o_pow(mf, (5.0 / 3.0))), rho2, o_fma(((7.0 / 3.0) * o_pow(mf, (4.0 / 3.0))), rho1, ((((((o_fma(chi2l, o_fma((-44.0), eta, o_fma((-81.0), delta, 81.0)), (chi1l * o_fma(81.0, (1.0 + delta), (-(44.0 * eta))))) * lalpi) / 48.0) + ((o_fma(1804.0, eta, (-969.0)) * o_pow(lalpi, (2.0 / 3.0))) / (1008.0 * o_pow(mf, (1.0 / 3.0))))) + (((o_fma(35371056.0, eta2, o_fma((24.0 * o_fma(857304.0, chi2l2, o_fma((-(994896.0 * chi1l)), chi2l, o_fma(857304.0, chi1l2, (-1975055.0))))), eta, o_fma((-(10287648.0 * chi1l2)), (1.0 + delta), o_fma((10287648.0 * chi2l2), delta, o_fma((-10287648.0), chi2l2, (-27312085.0)))))) * o_pow(lalpi, (4.0 / 3.0))) * o_pow(mf, (1.0 / 3.0))) / 6096384.0)) + ((((5.0 * o_pow(lalpi, (5.0 / 3.0))) * o_fma((42840.0 * o_fma(4.0, eta, (-1.0))), lalpi, o_fma(chi2l, o_fma((-35632.0), eta2, o_fma((-o_fma(6048.0, chi2l2, 287213.0)), ((-1.0) + delta), ((4.0 * o_fma(2083.0, delta, o_fma((1512.0 * chi2l2), ((-3.0) + delta), (-93414.0)))) * eta))), o_fma(((-6048.0) * chi1l3), o_fma((3.0 + delta), eta, ((-1.0) - delta)), (chi1l * o_fma((-35632.0), eta2, o_fma(287213.0, (1.0 + delta), (-((4.0 * o_fma(2083.0, delta, 93414.0)) * eta))))))))) * o_pow(mf, (2.0 / 3.0))) / 96768.0)) - ((((lalpi * lalpi) * o_fma((12.0 * eta), o_fma(21384760320.0, (lalpi * lalpi), o_fma((-(8345272320.0 * (chi1l + chi2l))), lalpi, o_fma((77616.0 * chi1l2), o_fma(282718.0, delta, (-158633.0)), o_fma((-(21943440288.0 * chi2l2)), delta, o_fma((-12312458928.0), chi2l2, o_fma((49568837472.0 * chi1l), chi2l, (-545384828789.0))))))), o_fma(7.0, o_fma((12160253952.0 * o_fma((-chi2l), delta, o_fma(chi1l, delta, (chi1l + chi2l)))), lalpi, o_fma((29362199328.0 * chi1l2), (1.0 + delta), o_fma((-(29362199328.0 * chi2l2)), delta, o_fma(29362199328.0, chi2l2, 177520268561.0)))), o_fma(((-336.0) * o_fma(1809550512.0, chi2l2, o_fma((-(2954929824.0 * chi1l)), chi2l, o_fma(1809550512.0, chi1l2, (-3248849057.0))))), eta2, (-(324322727232.0 * eta3)))))) * mf) / 30042980352.0)))));
}

// inspiral.py:141
template <class R> static inline R ins_phase_22_v3(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R eta4 = (eta3 * eta);
    R s2 = (s * s);
    R s3 = (s2 * s);
    R dchi2 = (dchi * dchi);
    R no_spin = (o_fma(8913612.508054944, eta4, o_fma((-3971998.0569125614), eta3, o_fma(376665.64637025696, eta2, o_fma(873401.6255736464, eta, 15415.000000000007)))) / o_fma(46.83697749859996, eta, 1.0));
    R eq_spin = ((s * o_fma(eta2, o_fma(895372.155565861, s2, o_fma(208349.45742548333, s, (-1982632.3844247842))), o_fma((-130668.37221912303), s2, o_fma(eta, o_fma((-174952.69161683554), s2, o_fma(1235279.439281927, s, (-1005307.3129700898))), o_fma(eta3, o_fma((-1689418.9124921719), s2, o_fma((-584728.050612325), s, 4662143.741417853)), o_fma((-207180.42746987), s, 397951.95299014193)))))) / o_fma(1.0, s3, o_fma(2.5298346636273306, s2, o_fma(3.5804521763363075, s, (-9.675704197652225)))));
    R uneq_spin = o_fma(((-1296.9289110696955) * dchi2), eta, (((dchi * delta) * eta) * o_fma(47752.17032707405, s, o_fma(24703.28267342699, eta, (-24708.109411857182)))));
    return ((no_spin + eq_spin) + uneq_spin);
}

// inspiral.py:155
template <class R> static inline R ins_phase_22_d13(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R s2 = (s * s);
    R s3 = (s2 * s);
    R s4 = (s3 * s);
    R no_spin = (o_fma(483033.0998073767, eta2, o_fma((-19943.076428555978), eta, (-17294.000000000007))) / o_fma(4.460294035404433, eta, 1.0));
    R eq_spin = ((s * o_fma(32614.091002011017, s4, o_fma(eta2, o_fma(1076587.7367729193, s3, o_fma((-137991.37032619823), s2, o_fma((-758386.5685734496), s, (-148368.4954044637)))), o_fma(19703.894135534803, s3, o_fma(eta, o_fma((-303141.1945565486), s3, o_fma(18298.307770807573, s2, o_fma(62190.404951852535, s, (-58475.33302037833)))), o_fma((-2179.3505885609297), s2, o_fma(67663.42759836042, s, 68384.62786426462))))))) / o_fma(1.0, s, 2.0412979553629143));
    R uneq_spin = (((12017.062595934838 * dchi) * delta) * eta);
    return ((no_spin + eq_spin) + uneq_spin);
}

// inspiral.py:167
template <class R> static inline R ins_phase_22_d23(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R s2 = (s * s);
    R s3 = (s2 * s);
    R no_spin = (o_fma((-557253.0066989232), eta3, o_fma(1169435.6931282217, eta2, o_fma((-120297.86185566607), eta, (-7579.300000000004)))) / o_fma(18.53018618227582, eta, 1.0));
    R eq_spin = ((s * o_fma(eta3, o_fma(847483.2999508682, s2, o_fma((-325423.99477314285), s, (-593633.5370110178))), o_fma(eta, o_fma(126736.43159783827, s2, o_fma(157036.45676788126, s, 50644.13475990821)), o_fma((-44331.41741405198), s2, o_fma(eta2, o_fma((-399712.22891153296), s2, o_fma((-50166.382087278434), s, 150022.21343386435)), o_fma((-66228.9369155027), s, (-27089.36915061857))))))) / o_fma(1.0, s3, o_fma((-1.130185486082531), s2, o_fma((-3.062957826830017), s, (-1.5232497464826662)))));
    R uneq_spin = (((3843.083992827935 * dchi) * delta) * eta);
    return ((no_spin + eq_spin) + uneq_spin);
}

// inspiral.py:179
template <class R> static inline R ins_phase_22_d43(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R s2 = (s * s);
    R no_spin = (o_fma(28867.73328134167, eta2, o_fma((-31133.52170083207), eta, 2439.000000000001)) / o_fma(0.41143032589262585, eta, 1.0));
    R eq_spin = ((s * o_fma(eta2, o_fma(117410.37432997991, s, 229284.04542668918), o_fma(9861.635308837876, s, o_fma(eta, o_fma((-25843.06175439942), s, (-82355.86732027541)), o_fma(eta3, o_fma((-386247.80765802023), s, (-375818.0132734753)), 16116.057657391262))))) / o_fma(1.0, s2, o_fma(0.25294420589064653, s, (-3.7385208695213668))));
    R uneq_spin = (((194.5554531509207 * dchi) * delta) * eta);
    return ((no_spin + eq_spin) + uneq_spin);
}

// inspiral.py:190
static inline double ins_inspiral_dphase_norm() {
    return (5.0 / (128.0 * lalpi().five_thirds));
}

// inspiral.py:195
template <class R> static inline R ins_phase_22_ansatz_exact(R mf, const Powers<R>& powers, const PhaseCoeffs<R>& phase) {
    R phase_in = phase.dphi0;
    phase_in = o_fma(phase.dphi2, powers.two_thirds, phase_in);
    phase_in = o_fma(phase.dphi3, mf, phase_in);
    phase_in = o_fma(phase.dphi4, powers.four_thirds, phase_in);
    phase_in = o_fma(phase.dphi5, powers.five_thirds, phase_in);
    phase_in = o_fma(phase.dphi6, powers.two, phase_in);
    phase_in = o_fma((phase.dphi6l * powers.two), powers.log, phase_in);
    phase_in = o_fma(phase.dphi7, powers.seven_thirds, phase_in);
    phase_in = o_fma(phase.dphi8, powers.eight_thirds, phase_in);
    phase_in = o_fma((phase.dphi8l * powers.eight_thirds), powers.log, phase_in);
    R pseudo = o_fma(phase.a0, powers.eight_thirds, (phase.a1 * powers.three));
    pseudo = o_fma((phase.a2 * powers.eight_thirds), powers.two_thirds, pseudo);
    pseudo = o_fma((phase.a3 * powers.eight_thirds), powers.itself, pseudo);
    phase_in = (phase_in + pseudo);
    return ((phase_in * powers.m_eight_thirds) * phase.dphase_norm);
}

// inspiral.py:214
template <class R> static inline R ins_phase_22_ansatz_int_exact(R mf, const Powers<R>& powers, const PhaseCoeffs<R>& phase) {
    R phasing = phase.phi0;
    phasing = o_fma(phase.phi2, powers.two_thirds, phasing);
    phasing = o_fma(phase.phi3, mf, phasing);
    phasing = o_fma(phase.phi4, powers.four_thirds, phasing);
    phasing = o_fma((phase.phi5l * powers.five_thirds), powers.log, phasing);
    phasing = o_fma(phase.phi6, powers.two, phasing);
    phasing = o_fma((phase.phi6l * powers.two), powers.log, phasing);
    phasing = o_fma(phase.phi7, powers.seven_thirds, phasing);
    phasing = o_fma(phase.phi8, powers.eight_thirds, phasing);
    phasing = o_fma((phase.phi8l * powers.eight_thirds), powers.log, phasing);
    R pseudo = o_fma(phase.sigma1, powers.eight_thirds, (phase.sigma2 * powers.three));
    pseudo = o_fma((phase.sigma3 * powers.one_third), powers.three, pseudo);
    pseudo = o_fma((phase.sigma4 * powers.two_thirds), powers.three, pseudo);
    phasing = (phasing + pseudo);
    return ((phasing * phase.phi_norm) * powers.m_five_thirds);
}

// ---- intermediate.py ----
// intermediate.py:16
template <class R> static inline R itm_amp_22_va(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R s2 = (s * s);
    R no_spin = (o_fma((-19837.908020966777), eta3, o_fma(27563.641024162127, eta2, o_fma(1974.6112656679577, eta, 1.4873184918202145))) / o_fma(458.4097306093354, eta2, o_fma(143.29004876335128, eta, 1.0)));
    R eq_spin = ((s * o_fma(eta3, o_fma(1486.4658089990298, s2, o_fma((-6962.675551371755), s, 1612.2681322644232)), o_fma((-19.38970173389662), s2, o_fma((3011.446602208493 * eta2), s, o_fma(3.2646808851249016, s, o_fma(eta, o_fma((-260.3494489873286), s, (-365.55631765202895)), 27.952730865904343)))))) / o_fma(1.0, s2, o_fma((-10.540154508599963), s, 12.647425554323242)));
    R uneq_spin = (((dchi * delta) * o_fma((-296.473359655246), eta, (-0.016404056649860943))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// intermediate.py:27
template <class R> static inline R itm_amp_22_delta0(R d1, R d4, R v1, R v2, R v4, R f1, R f2, R f4) {
    R f12 = (f1 * f1);
    R f42 = (f4 * f4);
    R f1mf2 = (f1 - f2);
    R f1mf4 = (f1 - f4);
    R f2mf4 = (f2 - f4);
    R f1mf22 = (f1mf2 * f1mf2);
    R f2mf42 = (f2mf4 * f2mf4);
    R f1mf43 = ((f1mf4 * f1mf4) * f1mf4);
    return (o_fma((((f12 * f1mf22) * f2) * o_fma(4.0, f42, o_fma((-(3.0 * f2)), f4, o_fma(f1, f2, (-((2.0 * f1) * f4)))))), v4, o_fma(f42, o_fma(((f2 * f2mf42) * o_fma((-f2), f4, o_fma((2.0 * f1), f4, o_fma((-4.0), f12, ((3.0 * f1) * f2))))), v1, ((f12 * f1mf43) * v2)), o_fma((-(((((d4 * f12) * f1mf22) * f1mf4) * f2) * f2mf4)), f4, ((((((d1 * f1) * f1mf2) * f1mf4) * f2) * f2mf42) * f42)))) / ((f1mf22 * f1mf43) * f2mf42));
}

// intermediate.py:40
template <class R> static inline R itm_amp_22_delta1(R d1, R d4, R v1, R v2, R v4, R f1, R f2, R f4) {
    R f12 = (f1 * f1);
    R f13 = (f12 * f1);
    R f14 = (f13 * f1);
    R f22 = (f2 * f2);
    R f23 = (f22 * f2);
    R f24 = (f23 * f2);
    R f42 = (f4 * f4);
    R f43 = (f42 * f4);
    R f44 = (f43 * f4);
    R f1mf2 = (f1 - f2);
    R f1mf4 = (f1 - f4);
    R f2mf4 = (f2 - f4);
    R f1mf22 = (f1mf2 * f1mf2);
    R f2mf42 = (f2mf4 * f2mf4);
    R f1mf43 = ((f1mf4 * f1mf4) * f1mf4);
    return (o_fma(((((d4 * f1) * f1mf22) * f1mf4) * f2mf4), o_fma((2.0 * f2), f4, (f1 * (f2 + f4))), (f4 * o_fma((-(((d1 * f1mf2) * f1mf4) * f2mf42)), o_fma((2.0 * f1), f2, ((f1 + f2) * f4)), (-((2.0 * f1) * o_fma(f1, o_fma((4.0 * f23), ((-v1) + v4), o_fma((2.0 * f43), ((-v1) + v2), (((6.0 * f22) * f4) * (v1 - v4)))), o_fma(((2.0 * f13) * f4), ((-v2) + v4), o_fma(((4.0 * f23) * f4), ((-v1) + v4), o_fma(f14, (v2 - v4), o_fma(f44, (v1 - v2), ((3.0 * f24) * (v1 - v4)))))))))))) / ((f1mf22 * f1mf43) * f2mf42));
}

// intermediate.py:60
template <class R> static inline R itm_amp_22_delta2(R d1, R d4, R v1, R v2, R v4, R f1, R f2, R f4) {
    R f12 = (f1 * f1);
    R f13 = (f12 * f1);
    R f14 = (f13 * f1);
    R f15 = (f14 * f1);
    R f22 = (f2 * f2);
    R f23 = (f22 * f2);
    R f24 = (f23 * f2);
    R f42 = (f4 * f4);
    R f43 = (f42 * f4);
    R f44 = (f43 * f4);
    R f45 = (f44 * f4);
    R f1mf2 = (f1 - f2);
    R f1mf4 = (f1 - f4);
    R f2mf4 = (f2 - f4);
    R f1mf22 = (f1mf2 * f1mf2);
    R f2mf42 = (f2mf4 * f2mf4);
    R f1mf43 = ((f1mf4 * f1mf4) * f1mf4);
    return (o_fma((-(f1mf22 * o_fma((f1 * o_fma(3.0, f2, (-(4.0 * f4)))), o_fma(2.0, f4, f2), o_fma(f12, o_fma(2.0, f2, f4), o_fma((f2 * o_fma(3.0, f2, (-(4.0 * f4)))), f4, f13))))), v4, o_fma((-f45), v2, o_fma((-(f1 * f44)), v2, o_fma(((8.0 * f12) * f43), v2, o_fma((-((8.0 * f13) * f42)), v2, o_fma((f14 * f4), v2, o_fma(f15, v2, o_fma(f45, v1, o_fma((f1 * f44), v1, o_fma((-((8.0 * f12) * f43)), v1, o_fma((-((4.0 * f23) * f42)), v1, o_fma((((12.0 * f12) * f2) * f42), v1, o_fma(((3.0 * f24) * f4), v1, o_fma((-(((4.0 * f1) * f23) * f4)), v1, o_fma(((3.0 * f1) * f24), v1, o_fma((-((4.0 * f12) * f23)), v1, o_fma((-(((d4 * f1mf22) * f1mf4) * f2mf4)), o_fma((2.0 * f1), (f2 + f4), o_fma(f2, f4, f12)), ((((d1 * f1mf2) * f1mf4) * f2mf42) * (o_fma(f1, f2, ((2.0 * (f1 + f2)) * f4)) + f42))))))))))))))))))) / ((f1mf22 * f1mf43) * f2mf42));
}

// intermediate.py:82
template <class R> static inline R itm_amp_22_delta3(R d1, R d4, R v1, R v2, R v4, R f1, R f2, R f4) {
    R f12 = (f1 * f1);
    R f13 = (f12 * f1);
    R f14 = (f13 * f1);
    R f22 = (f2 * f2);
    R f23 = (f22 * f2);
    R f24 = (f23 * f2);
    R f42 = (f4 * f4);
    R f43 = (f42 * f4);
    R f44 = (f43 * f4);
    R f1mf2 = (f1 - f2);
    R f1mf4 = (f1 - f4);
    R f2mf4 = (f2 - f4);
    R f1mf22 = (f1mf2 * f1mf2);
    R f2mf42 = (f2mf4 * f2mf4);
    R f1mf43 = ((f1mf4 * f1mf4) * f1mf4);
    return (o_fma(2.0, o_fma(((2.0 * f1) * f4), o_fma(((2.0 * f2) * f4), ((-v1) + v4), o_fma(f42, (v1 - v2), (f22 * (v1 - v4)))), o_fma(f14, ((-v2) + v4), o_fma(f24, ((-v1) + v4), o_fma(((2.0 * f13) * f4), (v2 - v4), o_fma(((2.0 * f22) * f42), (v1 - v4), o_fma(f44, ((-v1) + v2), (((2.0 * f12) * f2mf42) * (v1 - v4)))))))), o_fma((((d4 * f1mf22) * f1mf4) * f2mf4), (o_fma(2.0, f1, f2) + f4), (-((((d1 * f1mf2) * f1mf4) * f2mf42) * o_fma(2.0, f4, (f1 + f2)))))) / ((f1mf22 * f1mf43) * f2mf42));
}

// intermediate.py:102
template <class R> static inline R itm_amp_22_delta4(R d1, R d4, R v1, R v2, R v4, R f1, R f2, R f4) {
    R f12 = (f1 * f1);
    R f13 = (f12 * f1);
    R f22 = (f2 * f2);
    R f23 = (f22 * f2);
    R f42 = (f4 * f4);
    R f43 = (f42 * f4);
    R f1mf2 = (f1 - f2);
    R f1mf4 = (f1 - f4);
    R f2mf4 = (f2 - f4);
    R f1mf22 = (f1mf2 * f1mf2);
    R f2mf42 = (f2mf4 * f2mf4);
    R f1mf43 = ((f1mf4 * f1mf4) * f1mf4);
    return (o_fma((-(f1mf22 * o_fma((-3.0), f4, o_fma(2.0, f2, f1)))), v4, o_fma((-f43), v2, o_fma(((3.0 * f1) * f42), v2, o_fma((-((3.0 * f12) * f4)), v2, o_fma(f13, v2, o_fma(f43, v1, o_fma((-((3.0 * f1) * f42)), v1, o_fma((-((3.0 * f22) * f4)), v1, o_fma((((6.0 * f1) * f2) * f4), v1, o_fma((2.0 * f23), v1, o_fma((-((3.0 * f1) * f22)), v1, o_fma((-((d4 * f1mf22) * f1mf4)), f2mf4, (((d1 * f1mf2) * f1mf4) * f2mf42))))))))))))) / ((f1mf22 * f1mf43) * f2mf42));
}

// intermediate.py:124
template <class R> static inline R itm_amp_22_ansatz(R mf, const Powers<R>& powers, const AmpCoeffs<R>& amp) {
    R polynomial = (amp.delta0 + (mf * (amp.delta1 + (mf * (amp.delta2 + (mf * (amp.delta3 + (mf * amp.delta4))))))));
    return (powers.seven_sixths / polynomial);
}

// intermediate.py:130
template <class R> static inline R itm_phase_22_v2(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R s2 = (s * s);
    R s3 = (s2 * s);
    R s4 = (s3 * s);
    R no_spin = (o_fma(1466725833.4378073, eta3, o_fma((-352257424.21184325), eta2, o_fma((-5581973.49185435), eta, (-82.54500000000004)))) / o_fma(2517658.5751772933, eta3, o_fma(5385164.380400193, eta2, o_fma(66757.12830903867, eta, 1.0))));
    R eq_spin = ((s * o_fma(eta, o_fma(37.462217031512786, s4, o_fma(29.071788188638315, s3, o_fma((-17.671713040498304), s2, o_fma((-132.88417400679026), s, 207.69898051583655)))), o_fma(4.984750041013893, s4, o_fma(eta3, o_fma((-866.283566780576), s4, o_fma(1897.5444343138167, s3, o_fma(415.7134909564443, s2, o_fma(1152.425940764218, s, (-1365.1499998427248))))), o_fma(5.95010003393006, s3, o_fma(eta2, o_fma((-647.8103976942541), s3, o_fma((-107.41099349364234), s, 170.97203068800542)), o_fma((-0.8612656616290079), s2, o_fma((-36.066611959079935), s, 19.416719811164853)))))))) / o_fma(1.0, s, (-1.1492259468169692)));
    R uneq_spin = (((dchi * delta) * eta3) * o_fma(515.9898508588834, s, o_fma((-20486.813161100774), eta, 7343.130973149263)));
    return ((no_spin + eq_spin) + uneq_spin);
}

// intermediate.py:143
template <class R> static inline R itm_phase_22_v2m_rdv4(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R s2 = (s * s);
    R no_spin = ((eta * o_fma(632.4731389009143, eta2, o_fma(101.21991715215253, eta, 0.9951733419499662))) / o_fma(1.0, eta3, o_fma(1.8413983770369362, eta2, o_fma(0.11412314719189287, eta, 0.00016803066316882238))));
    R eq_spin = ((s * o_fma(eta3, o_fma(178.85850322465944, s2, o_fma((-14302.70838220423), s, 3531.552143264721)), o_fma(0.3612417066833153, s2, o_fma(eta, o_fma((-147.53381808989846), s2, (-697.6773920613674)), o_fma((4941.31613710257 * eta2), s, o_fma(16.89845522539974, s, 18.694178521101332)))))) / o_fma(1.0, s2, o_fma((-2.7706595614504725), s, 2.965640445745779)));
    R uneq_spin = (((dchi * delta) * eta2) * o_fma((1693.326644293169 * eta2), s, 356.74395864902294));
    return ((no_spin + eq_spin) + uneq_spin);
}

// intermediate.py:154
template <class R> static inline R itm_phase_22_v3m_rdv4(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R eta6 = (eta3 * eta3);
    R s2 = (s * s);
    R no_spin = ((eta * o_fma((-751.4184178636324), eta3, o_fma(688.3609087244353, eta2, o_fma((-227.46830225846668), eta, (-5.126358906504587))))) / o_fma(1.0, eta2, o_fma((-0.7811680872741462), eta, (-0.004551938711031158))));
    R eq_spin = ((s * o_fma(eta3, o_fma(224.77581754671272, s2, o_fma(1862.9097047992134, s, (-936.3740515136005))), o_fma((-2.84311102369862), s2, o_fma(eta, o_fma((-8.13494176717772), s2, 73.79645135116367), o_fma((-(539.4071941841604 * eta2)), s, o_fma((-0.9539250460041732), s, 0.1549280856660919)))))) / o_fma(1.0, s, (-1.5308507364054487)));
    R uneq_spin = (((2993.3598520496153 * dchi) * delta) * eta6);
    return ((no_spin + eq_spin) + uneq_spin);
}

// intermediate.py:166
template <class R> static inline R itm_phase_22_d43(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R eta4 = (eta3 * eta);
    R s2 = (s * s);
    R no_spin = (o_fma(670077.5414916876, eta4, o_fma((-967049.2793750163), eta3, o_fma((-282820.39946006844), eta2, o_fma((-906.746595921514), eta, 0.4248820426833804)))) / o_fma(19783.077247023448, eta2, o_fma(1670.9440812294847, eta, 1.0)));
    R eq_spin = ((s * o_fma(eta4, o_fma(8745.201037897836, s2, o_fma((-227.7672020783411), s, (-4939.642457025497))), o_fma(eta2, o_fma(941.6974723887743, s2, o_fma(269.9234918621958, s, (-711.8532052499075))), o_fma(0.4818323187946999, s2, o_fma(eta, o_fma((-61.17248283184154), s2, 12.840649528989287), o_fma(eta3, o_fma((-4974.189172654984), s2, o_fma((-877.8811492839261), s, 3499.432393555856)), o_fma(1.1366593671801855, s, 0.22814271667259703))))))) / o_fma(1.0, s, (-1.2442293719740283)));
    R uneq_spin = (((dchi * delta) * o_fma(1493.3851099678195, eta, (-514.8494071830514))) * eta3);
    return ((no_spin + eq_spin) + uneq_spin);
}

// intermediate.py:178
template <class R> static inline R itm_phase_22_ansatz_exact(R mf, const Powers<R>& powers, const PhaseCoeffs<R>& phase) {
    R fda = phase.f_damp;
    R dfr = (mf - phase.f_ring);
    R lorentzian = ((4.0 * phase.c_l_gr) / o_fma((4.0 * fda), fda, (dfr * dfr)));
    R phase_out = o_fma(phase.b1, powers.m_one, phase.b0);
    phase_out = o_fma(phase.b2, powers.m_two, phase_out);
    phase_out = o_fma(phase.b3, powers.m_three, phase_out);
    phase_out = o_fma(phase.b4, powers.m_four, phase_out);
    return (phase_out + lorentzian);
}

// intermediate.py:190
template <class R> static inline R itm_phase_22_ansatz_int_exact(R mf, const Powers<R>& powers, const PhaseCoeffs<R>& phase) {
    R fda = phase.f_damp;
    R phase_out = o_fma(phase.b0, mf, (phase.b1 * powers.log));
    phase_out = o_fma((-phase.b2), powers.m_one, phase_out);
    phase_out = (phase_out - ((phase.b3 * powers.m_two) / 2.0));
    phase_out = (phase_out - ((phase.b4 * powers.m_three) / 3.0));
    return (phase_out + (((2.0 * phase.c_l_gr) * o_atan(((mf - phase.f_ring) / (2.0 * fda)))) / fda));
}

// ---- ringdown.py ----
// ringdown.py:16
template <class R> static inline R rd_amp_22_gamma2(R eta, R s, R dchi, R delta) {
    R s2 = (s * s);
    R eta2 = (eta * eta);
    R no_spin = (o_fma((-18.256121237800397), eta2, o_fma(7.480371544268765, eta, 0.8312293675316895)) / o_fma((-30.578409433912874), eta2, o_fma(10.915453595496611, eta, 1.0)));
    R eq_spin = ((s * o_fma(eta2, o_fma(4.6849496672664594, s, 0.25295441250444334), o_fma(0.031852563636196894, s, o_fma(eta, o_fma((-2.8489481072076472), s, (-0.1467158405070222)), 0.5869408584532747)))) / o_fma(1.0, s2, o_fma((-3.41755361841226), s, 3.8775263105069953)));
    R uneq_spin = ((((-0.00548054788508203) * dchi) * delta) * eta);
    return ((no_spin + eq_spin) + uneq_spin);
}

// ringdown.py:26
template <class R> static inline R rd_amp_22_gamma3(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R no_spin = (o_fma((-4.222259944408823), eta3, o_fma(2.109081209912545, eta2, o_fma((-4.091333144596439), eta, 1.3666000000000007))) / o_fma((-2.7440263888207594), eta, 1.0));
    R eq_spin = (o_fma((-0.05633734476062242), s, o_fma(eta, o_fma(0.4168560229353532, s, (-0.8752427297525086)), o_fma(eta2, o_fma((-0.6330998412809531), s, 2.331724812782498), 0.07179105336478316))) * s);
    R uneq_spin = ((0.0 * delta) * dchi);
    return ((no_spin + eq_spin) + uneq_spin);
}

// ringdown.py:36
template <class R> static inline R rd_amp_22_v1(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R s2 = (s * s);
    R s3 = (s * s2);
    R s4 = (s * s3);
    R no_spin = (o_fma(162.52904393600332, eta2, o_fma(25.417967754401182, eta, 0.03689164742964719)) / o_fma((-29.628854485544874), eta2, o_fma(61.19874463331437, eta, 1.0)));
    R eq_spin = ((s * o_fma(eta2, o_fma(72.08174136362386, s4, o_fma((-5.104257870393138), s3, o_fma((-47.99096500250743), s2, o_fma((-2.398327419614959), s, 27.675454081988036)))), o_fma((-0.18147275151697131), s4, o_fma(eta, o_fma((-8.990840289951514), s4, o_fma((-0.2874540719094058), s3, o_fma(6.176053843938542, s2, o_fma(1.11147906765112, s, (-4.805034453745424))))), o_fma(eta3, o_fma((-132.77982622925845), s4, o_fma(30.491948143930266, s3, o_fma(81.9290740950083, s2, o_fma((-3.751501972663298), s, (-48.31945248941757))))), o_fma((-0.05292913111731128), s3, o_fma(0.19967405175523437, s2, o_fma(0.026356911108320547, s, (-0.14352506969368556))))))))) / o_fma(1.0, s, (-1.4160870461211452)));
    R uneq_spin = ((((-0.04426571511345366) * dchi) * delta) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// ringdown.py:49
template <class R> static inline R rd_amp_22_ansatz(R mf, const AmpCoeffs<R>& amp) {
    R dfr = (mf - amp.f_ring);
    return ((o_exp(((-dfr) * amp.gamma_r)) * amp.gamma_d13) / ((dfr * dfr) + amp.gamma_d2));
}

// ringdown.py:55
template <class R> static inline R rd_amp_22_ansatz_exact(R mf, R f_ring, R f_damp, const AmpCoeffs<R>& amp) {
    R gamma_d13 = ((f_damp * amp.gamma1) * amp.gamma3);
    R gamma_r = (amp.gamma2 / (amp.gamma3 * f_damp));
    R gamma_d2 = ((f_damp * amp.gamma3) * (f_damp * amp.gamma3));
    R dfr = (mf - f_ring);
    return ((o_exp(((-dfr) * gamma_r)) * gamma_d13) / o_fma(dfr, dfr, gamma_d2));
}

// ringdown.py:64
template <class R> static inline R rd_amp_22_dansatz(R mf, R f_ring, R f_damp, const AmpCoeffs<R>& amp) {
    R g1 = amp.gamma1;
    R g2 = amp.gamma2;
    R g3 = amp.gamma3;
    R frd = f_ring;
    R fda = f_damp;
    R dfr = (mf - frd);
    R dfd = (fda * g3);
    R prefactor = ((-o_exp((((-g2) * dfr) / dfd))) * g1);
    R numerator = o_fma((((fda * fda) * g2) * g3), g3, o_fma((dfr * dfr), g2, (((2.0 * fda) * dfr) * g3)));
    R denominator = (o_fma(dfr, dfr, (dfd * dfd)) * o_fma(dfr, dfr, (dfd * dfd)));
    return ((prefactor * numerator) / denominator);
}

// ringdown.py:79
template <class R> static inline R rd_amp_22_peak_frequency(R gamma2, R gamma3, R frd, R fda) {
    return o_where((gamma2 <= 1.0), o_abs((frd + (((fda * gamma3) * (o_sqrt_nonneg(o_fma((-gamma2), gamma2, 1.0)) - 1.0)) / gamma2))), o_abs((frd + (((fda * (-1.0)) * gamma3) / gamma2))));
}

// ringdown.py:84
template <class R> static inline R rd_phase_22_v4(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R eta4 = (eta3 * eta);
    R eta5 = (eta4 * eta);
    R s2 = (s * s);
    R s3 = (s2 * s);
    R s4 = (s3 * s);
    R no_spin = (o_fma((-39783.51698102803), eta5, o_fma(12828.269960300782, eta4, o_fma(7732.064464348168, eta3, o_fma((-4925.756920247186), eta2, o_fma((-4616.740713893726), eta, (-85.86062966719405)))))) / o_fma(50.206318806624004, eta, 1.0));
    R eq_spin = ((s * o_fma(eta4, o_fma(5064.173605639933, s4, o_fma((-608.2901586790335), s3, o_fma((-4253.974688619423), s2, o_fma((-3443.0233614187396), s, 1075.8686153198323)))), o_fma(eta, o_fma(97.69027029734173, s4, o_fma(31.231253209893488, s3, o_fma((-15.399770764623746), s2, o_fma(26.580344399838758, s, (-69.19412903018717))))), o_fma(1.6533417657003922, s4, o_fma(eta2, o_fma((-719.8775484010988), s4, o_fma((-104.36120236420928), s3, o_fma(423.48863373726243, s2, o_fma((-18.184492163348665), s, 93.64156367505917)))), o_fma(2.302712009652155, s3, o_fma((-3.835967351280833), s2, o_fma((eta3 * o_fma((-101.72731770500685), s, 1497.3545918387515)), s, o_fma((-36.49019206094966), s, 33.335857451144356))))))))) / o_fma(1.0, s, (-1.3705601055555852)));
    R uneq_spin = (((dchi * delta) * eta) * o_fma(156.08206945239374, eta, 22.363215261437862));
    return ((no_spin + eq_spin) + uneq_spin);
}

// ringdown.py:99
template <class R> static inline R rd_phase_22_d12(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R eta4 = (eta3 * eta);
    R s2 = (s * s);
    R s3 = (s2 * s);
    R s4 = (s3 * s);
    R s5 = (s4 * s);
    R no_spin = ((eta * o_fma(6.086871214811216, eta2, o_fma((-1.237332073800276), eta, 0.7207992174994245))) / o_fma(1.0, eta3, o_fma((-0.15500218299268662), eta2, o_fma(0.06099184229137391, eta, 0.006851189888541745))));
    R eq_spin = (o_fma(o_fma((-11.61488280763592), eta2, o_fma((-22.384949087140086), eta4, o_fma(4.917070939324979, eta, (-0.2697933899920511)))), s5, o_fma((eta * o_fma(338.7263666984089, eta3, o_fma((-43.97332874253772), eta, 5.7756520242745735))), s4, o_fma(o_fma((-47.56042058800358), eta4, o_fma((-0.28745205203100666), eta, 0.20035146870472367)), s3, o_fma(o_fma(160.14971486043524, eta3, o_fma(58.59408241189781, eta2, o_fma((-308.62513664956975), eta4, o_fma((-25.25397971063995), eta, 0.06519048552628343)))), s, ((eta * o_fma(835.1725103648205, eta3, o_fma((-693.0504179144295), eta2, o_fma(153.95945758807616, eta, (-5.215945111216946))))) * s2))))) / o_fma((-0.6628745847248266), s, 1.0));
    R uneq_spin = ((((-23.504907495268824) * dchi) * delta) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// ringdown.py:114
template <class R> static inline R rd_phase_22_d24(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R eta4 = (eta3 * eta);
    R s2 = (s * s);
    R s3 = (s2 * s);
    R no_spin = ((eta * o_fma(64.69109972468395, eta2, o_fma(9.429314399633007, eta, (-9.460253118496386)))) / o_fma(1.0, eta2, o_fma((-0.09987544893382533), eta, (-0.0670554310666559))));
    R eq_spin = (o_fma(0.04497628581617564, s3, o_fma((eta2 * s), o_fma(64.42947340363101, s2, o_fma((-62.997389062600035), s, (-191.00932194869588))), o_fma((eta4 * s), o_fma((-1031.846477275069), s2, o_fma((-2177.554979351284), s, (-774.3633787391745))), o_fma((17.36495157980372 * eta), s, ((eta3 * s) * o_fma(808.457330742532, s, 930.3458437154668)))))) / o_fma((-0.7267610313751913), s, 1.0));
    R uneq_spin = (((dchi * delta) * o_fma(91.60477826830407, eta, (-36.66374091965371))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// ringdown.py:127
template <class R> static inline R rd_phase_22_d34(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R eta5 = (eta3 * eta2);
    R s2 = (s * s);
    R s3 = (s2 * s);
    R s4 = (s3 * s);
    R no_spin = ((eta * o_fma(13.936621412517798, eta, (-8.506898502692536))) / o_fma(1.0, eta, (-0.40919671232073945)));
    R eq_spin = (o_fma((eta3 * s), o_fma(109.04824706217418, s2, o_fma((-633.1966645925428), s3, o_fma(6.404060932334562, s, (-27.976989112929353)))), o_fma(0.046849371468156265, s2, o_fma(eta, o_fma((-13.743271480938104), s4, o_fma(1.7280582989361533, s, (18.41570325463385 * s3))), (eta2 * o_fma(215.78111099820157, s4, o_fma(73.8367329022058, s, (-(95.57802408341716 * s3)))))))) / o_fma((-0.6862449113932192), s, 1.0));
    R uneq_spin = (((641.8965762829259 * dchi) * delta) * eta5);
    return ((no_spin + eq_spin) + uneq_spin);
}

// ringdown.py:141
template <class R> static inline R rd_phase_22_d54(R eta, R s, R dchi, R delta) {
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R no_spin = ((eta * o_fma(119.43820622871043, eta2, o_fma(22.455288821807095, eta, 7.05731400277692))) / o_fma(1.0, eta, 0.26026709603623255));
    R eq_spin = (o_fma((eta3 * s), o_fma(90.31815139272628, s, (-316.26970506215554)), o_fma((eta2 * o_fma((-56.05992404859163), s, 134.88158268621922)), s, ((eta * s) * o_fma(9.486783128047414, s, (-7.9407123129681425))))) / o_fma((-0.7162058321905909), s, 1.0));
    R uneq_spin = (((43.82713604567481 * dchi) * delta) * eta3);
    return ((no_spin + eq_spin) + uneq_spin);
}

// ringdown.py:151
template <class R> static inline R rd_phase_22_ansatz_exact(R mf, const Powers<R>& powers, const PhaseCoeffs<R>& phase) {
    R fda = phase.f_damp;
    R dfr = (mf - phase.f_ring);
    R phase_rd = o_fma(phase.c1, powers.m_one_third, phase.c0);
    phase_rd = o_fma(phase.c2, powers.m_two, phase_rd);
    phase_rd = o_fma(phase.c4, powers.m_four, phase_rd);
    return (phase_rd + (phase.c_l / o_fma(fda, fda, (dfr * dfr))));
}

// ringdown.py:161
template <class R> static inline R rd_phase_22_ansatz_int_exact(R mf, const Powers<R>& powers, const PhaseCoeffs<R>& phase) {
    R phase_rd = o_fma(phase.c0, mf, ((1.5 * phase.c1) * powers.two_thirds));
    phase_rd = o_fma((-phase.c2), powers.m_one, phase_rd);
    phase_rd = o_fma((-phase.c4_ov3), powers.m_three, phase_rd);
    return o_fma(phase.c_l_ov_fda, o_atan(((mf - phase.f_ring) / phase.f_damp)), phase_rd);
}

// qnm.py:32: fring22 numerator, coefficients of af^0..af^7
static const double qnm_fring22_numerator[8] = {0.05947169566573468, -0.14989771215394762, 0.09535606290986028, 0.02260924869042963, -0.02501704155363241, -0.005852438240997211, 0.0027489038393367993, 0.0005821983163192694};
// qnm.py:33: fring22 denominator, coefficients of af^0..af^6
static const double qnm_fring22_denominator[7] = {1.0, -2.8570126619966296, 2.373335413978394, 0.0, -0.6036964688511505, 0.0, 0.0873798215084077};

// qnm.py:44: fdamp22 numerator, coefficients of af^0..af^6
static const double qnm_fdamp22_numerator[7] = {0.014158792290965177, -0.036989395871554566, 0.026822526296575368, 0.0008490933750566702, -0.004843996907020524, -0.00014745235759327472, 0.0001504546201236794};
// qnm.py:45: fdamp22 denominator, coefficients of af^0..af^6
static const double qnm_fdamp22_denominator[7] = {1.0, -2.5900842798681376, 1.8952576220623967, 0.0, -0.31416610693042507, 0.0, 0.009002719412204133};

// amp22.py:26 (amp_coefficients): the statements assigning pn_two_thirds, pn_three_thirds, pn_four_thirds, pn_five_thirds, pn_six_thirds
template <class R> static inline void amp22_pn_terms(R chi1l, R chi1l2, R chi1l3, R chi2l, R chi2l2, R delta, R eta, R eta2, R eta3, R& pn_two_thirds, R& pn_three_thirds, R& pn_four_thirds, R& pn_five_thirds, R& pn_six_thirds) {
    pn_two_thirds = ((o_fma(1804.0, eta, (-969.0)) / 672.0) * lalpi().two_thirds);
    pn_three_thirds = ((o_fma((-(44.0 * (chi1l + chi2l))), eta, o_fma((-(81.0 * chi2l)), delta, o_fma(81.0, (chi1l + chi2l), ((81.0 * chi1l) * delta)))) / 48.0) * lalpi().itself);
    pn_four_thirds = (((o_fma(24.0, o_fma(1473794.0, eta2, o_fma((428652.0 * chi2l2), ((-1.0) + delta), (o_fma(10584.0, o_fma(81.0, chi2l2, o_fma(81.0, chi1l2, (-((94.0 * chi1l) * chi2l)))), (-1975055.0)) * eta))), o_fma((-(10287648.0 * chi1l2)), (1.0 + delta), (-27312085.0))) / 8128512.0) * lalpi().one_third) * lalpi().itself);
    pn_five_thirds = ((o_fma((42840.0 * o_fma(4.0, eta, (-1.0))), M_PI_, o_fma(chi1l, o_fma(287213.0, (1.0 + delta), (-((4.0 * eta) * o_fma(8908.0, eta, o_fma(2083.0, delta, 93414.0))))), o_fma(((-6048.0) * chi1l3), o_fma((3.0 + delta), eta, ((-1.0) - delta)), (chi2l * o_fma((-35632.0), eta2, o_fma((-o_fma(6048.0, chi2l2, 287213.0)), ((-1.0) + delta), ((4.0 * o_fma(2083.0, delta, o_fma((1512.0 * chi2l2), ((-3.0) + delta), (-93414.0)))) * eta))))))) / 32256.0) * lalpi().five_thirds);
    pn_six_thirds = ((o_fma(12.0, o_fma(eta, o_fma((-77616.0), o_fma(275520.0, lalpi().two, o_fma((-(107520.0 * (chi1l + chi2l))), M_PI_, o_fma((-chi2l2), o_fma(282718.0, delta, 158633.0), o_fma((638642.0 * chi1l), chi2l, (chi1l2 * o_fma(282718.0, delta, (-158633.0))))))), 545384828789.0), o_fma((-116424.0), o_fma(147117.0, o_fma((-chi2l2), ((-1.0) + delta), (chi1l2 * (1.0 + delta))), ((60928.0 * o_fma((-chi2l), delta, o_fma(chi1l, delta, (chi1l + chi2l)))) * M_PI_)), o_fma((28.0 * o_fma(11088.0, o_fma(163199.0, chi2l2, o_fma(163199.0, chi1l2, (-((266498.0 * chi1l) * chi2l)))), (-3248849057.0))), eta2, (27026893936.0 * eta3)))), (-1242641879927.0)) / 60085960704.0) * lalpi().two);
}

// phase22.py:53 (phase_coefficients): the statements assigning phi2_ns, phi3_ns, phi3_s, phi4_ns, phi4_s, phi5l_ns, phi5l_s, phi6_ns, phi6_s, phi6l_ns, phi7_ns, phi7_s, phi8_s, phi8l_s
template <class R> static inline void phase22_pn_terms(R chi1l, R chi1l2, R chi1l2l, R chi1l3, R chi2l, R chi2l2, R chi2l3, R delta, R eta, R eta2, R eta3, R& phi2_ns, R& phi3_ns, R& phi3_s, R& phi4_ns, R& phi4_s, R& phi5l_ns, R& phi5l_s, R& phi6_ns, R& phi6_s, R& phi6l_ns, R& phi7_ns, R& phi7_s, R& phi8_s, R& phi8l_s) {
    phi2_ns = (((3715.0 / 756.0) + ((55.0 * eta) / 9.0)) * lalpi().two_thirds);
    phi3_ns = ((-16.0) * lalpi().two);
    phi3_s = ((o_fma(113.0, o_fma((-chi2l), delta, o_fma(chi1l, delta, (chi1l + chi2l))), (-((76.0 * (chi1l + chi2l)) * eta))) / 6.0) * lalpi().itself);
    phi4_ns = ((((15293365.0 / 508032.0) + ((27145.0 * eta) / 504.0)) + ((3085.0 * eta2) / 72.0)) * lalpi().four_thirds);
    phi4_s = ((((-5.0) * o_fma((-(81.0 * chi2l2)), o_fma(2.0, eta, ((-1.0) + delta)), o_fma((81.0 * chi1l2), o_fma((-2.0), eta, (1.0 + delta)), ((316.0 * chi1l2l) * eta)))) / 16.0) * lalpi().four_thirds);
    phi5l_ns = ((((5.0 * o_fma((-6552.0), eta, 46374.0)) * M_PI_) / 4536.0) * lalpi().five_thirds);
    phi5l_s = ((o_fma((85680.0 * (chi1l + chi2l)), eta2, o_fma((-732985.0), o_fma((-chi2l), delta, o_fma(chi1l, delta, (chi1l + chi2l))), (-((560.0 * o_fma((-1213.0), (chi1l + chi2l), ((63.0 * (chi1l - chi2l)) * delta))) * eta)))) / 4536.0) * lalpi().five_thirds);
    phi6_ns = ((((((((11583231236531.0 / 4694215680.0) - (((5.0 * eta) * o_fma((588.0 * eta), o_fma(102260.0, eta, (-45633.0)), 3147553127.0)) / 3048192.0)) - ((6848.0 * EGAMMA_) / 21.0)) - ((640.0 * lalpi().two) / 3.0)) + (((2255.0 * eta) * lalpi().two) / 12.0)) - ((13696.0 * LN2_) / 21.0)) - ((6848.0 * lalpi().log) / 63.0)) * lalpi().two);
    phi6_s = ((((5.0 * o_fma(227.0, o_fma((-chi2l), delta, o_fma(chi1l, delta, (chi1l + chi2l))), (-((156.0 * (chi1l + chi2l)) * eta)))) * M_PI_) / 3.0) * lalpi().two);
    phi6_s = o_fma(((5.0 * o_fma((-(7.0 * chi1l2)), o_fma(9808.0, eta2, o_fma((-15103.0), (1.0 + delta), ((2.0 * o_fma(6580.0, delta, 21683.0)) * eta))), o_fma(((20.0 * chi1l2l) * eta), o_fma(12488.0, eta, 11763.0), ((7.0 * chi2l2) * o_fma((-9808.0), eta2, o_fma((-15103.0), ((-1.0) + delta), ((2.0 * o_fma(6580.0, delta, (-21683.0))) * eta))))))) / 4032.0), lalpi().two, phi6_s);
    phi6l_ns = (((-6848.0) / 63.0) * lalpi().two);
    phi7_ns = ((((5.0 * o_fma((168.0 * o_fma((-29618.0), eta, 75703.0)), eta, 15419335.0)) * M_PI_) / 254016.0) * lalpi().seven_thirds);
    phi7_s = (((5.0 * o_fma((90514368.0 * (chi1l + chi2l)), eta3, o_fma((-(1008.0 * o_fma(208433.0, (chi1l + chi2l), ((25011.0 * (chi1l - chi2l)) * delta)))), eta2, o_fma((-5030016755.0), o_fma((-chi2l), delta, o_fma(chi1l, delta, (chi1l + chi2l))), ((4.0 * o_fma(2113331119.0, (chi1l + chi2l), ((675484362.0 * (chi1l - chi2l)) * delta))) * eta))))) / 6096384.0) * lalpi().seven_thirds);
    phi7_s = o_fma((((-5.0) * o_fma((-(57.0 * chi2l2)), o_fma(2.0, eta, ((-1.0) + delta)), o_fma((57.0 * chi1l2), o_fma((-2.0), eta, (1.0 + delta)), ((220.0 * chi1l2l) * eta)))) * M_PI_), lalpi().seven_thirds, phi7_s);
    phi7_s = o_fma((o_fma(((40.0 * (chi1l + chi2l)) * o_fma(17.0, chi2l2, o_fma(17.0, chi1l2, (-(14.0 * chi1l2l))))), eta2, o_fma(14585.0, o_fma((-chi2l3), ((-1.0) + delta), (chi1l3 * (1.0 + delta))), (-((5.0 * o_fma(chi1l3, o_fma(2985.0, delta, 8819.0), o_fma((-((8439.0 * chi1l2) * chi2l)), (1.0 + delta), o_fma(chi2l3, o_fma((-2985.0), delta, 8819.0), (((8439.0 * chi1l) * chi2l2) * ((-1.0) + delta)))))) * eta)))) / 48.0), lalpi().seven_thirds, phi7_s);
    R phi8_common = o_fma((94344.0 * (chi1l + chi2l)), eta2, o_fma(1263141.0, o_fma((-chi2l), delta, o_fma(chi1l, delta, (chi1l + chi2l))), (-((2.0 * o_fma(794075.0, (chi1l + chi2l), ((178533.0 * (chi1l - chi2l)) * delta))) * eta))));
    phi8_s = ((((((-5.0) * phi8_common) * M_PI_) * ((-1.0) + lalpi().log)) / 9072.0) * lalpi().eight_thirds);
    phi8l_s = (((((-5.0) * phi8_common) * M_PI_) / 9072.0) * lalpi().eight_thirds);
}

