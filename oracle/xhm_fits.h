// GENERATED by oracle/gen_fits.py from the reference's
// src/waveformx/models/xpnr/_generated/hm_*.py (LALSimulation 6.2.0's IMRPhenomXHM fits via
// waveformrs; GPL-3.0-or-later data). Oracle only: generic over the scalar R, fma(a, b, c)
// unfused as a * b + c. Do not edit.
#pragma once

// ---- hm_qnm.py: numerator and denominator in ascending powers of af ----
// hm_qnm.py:18
static const double hmq_fring21_num[6] = {0.059471695665734674, -0.07585416297991414, 0.021967909664591865, -0.0018964744613388146, 0.001164879406179587, -0.0003387374454044957};
static const double hmq_fring21_den[3] = {1.0, -1.4437415542456158, 0.49246920313191234};
// hm_qnm.py:27
static const double hmq_fring33_num[7] = {0.09540436245212061, -0.22799517865876945, 0.13402916709362475, 0.03343753057911253, -0.030848060170259615, -0.006756504382964637, 0.0027301732074159835};
static const double hmq_fring33_den[7] = {1.0, -2.7265947806178334, 2.144070539525238, 0.0, -0.4706873667569393, 0.0, 0.05321818246993958};
// hm_qnm.py:37
static const double hmq_fring32_num[7] = {0.09540436245212061, -0.13628306966373951, 0.030099881830507727, -0.000673589757007597, 0.0118277880067919, 0.0020533816327907334, -0.0015206141948469621};
static const double hmq_fring32_den[7] = {1.0, -1.6531854335715193, 0.5634705514193629, 0.0, 0.12256204148002939, 0.0, -0.027297817699401976};
// hm_qnm.py:47
static const double hmq_fring44_num[7] = {0.1287821193485683, -0.21224284094693793, 0.0710926778043916, 0.015487322972031054, -0.002795401084713644, 4.5483523029172406e-05, 0.00034775290179000503};
static const double hmq_fring44_den[5] = {1.0, -1.9931645124693607, 1.0593147376898773, 0.0, -0.06378640753152783};
// hm_qnm.py:57
static const double hmq_fdamp21_num[6] = {2.0696914454467294, -3.1358071947583093, 0.14456081596393977, 1.2194717985037946, -0.2947372598589144, 0.002943057145913646};
static const double hmq_fdamp21_den[5] = {146.1779212636481, -219.81790388304876, 17.7141194900164, 75.90115083917898, -18.975287709794745};
// hm_qnm.py:66
static const double hmq_fdamp33_num[6] = {0.014754148319335946, -0.03124423610028678, 0.017192623913708124, 0.001034954865629645, -0.0015925124814622795, -0.0001414350555699256};
static const double hmq_fdamp33_den[5] = {1.0, -2.0963684630756894, 1.196809702382645, 0.0, -0.09874113387889819};
// hm_qnm.py:75
static const double hmq_fdamp32_num[5] = {0.014754148319335946, -0.03445752346074498, 0.02168855041940869, 0.0014945908223317514, -0.0034761714223258693};
static const double hmq_fdamp32_den[5] = {1.0, -2.320722660848874, 1.5096146036915865, 0.0, -0.18791187563554512};
// hm_qnm.py:83
static const double hmq_fdamp44_num[7] = {0.014986847152355699, -0.01722587715950451, -0.0016734788189065538, 0.0002837322846047305, 0.002510528746148588, 0.00031983835498725354, 0.000812185411753066};
static const double hmq_fdamp44_den[7] = {1.0, -1.1350205970682399, -0.0500827971270845, 0.0, 0.13983808071522857, 0.0, 0.051876225199833995};
// hm_qnm.py:93
static const double hmq_re_l2m2lp2_num[6] = {1.0, -2.2956993576253635, 1.461988775298876, 0.0043296365593147035, -0.1695667458204109, -0.0006267849034466508};
static const double hmq_re_l2m2lp2_den[7] = {1.0, -2.2956977727459043, 1.4646339137818438, 0.0, -0.16843226886562457, 0.0, -7.150540890128118e-05};
// hm_qnm.py:103
static const double hmq_re_l2m2lp3_num[7] = {0.0, 18.522563276099167, -37.978140351289014, 19.030390708998894, 3.0355668591803386, -2.210028290847915, -0.37117112862247975};
static const double hmq_re_l2m2lp3_den[7] = {164.52480238697507, -377.9093045285145, 243.3353695550844, 0.0, -30.79738566181734, 0.0, 1.0};
// hm_qnm.py:113
static const double hmq_re_l3m2lp2_num[7] = {0.0, 0.47513455283841244, -0.9016636384605536, 0.3844811236426182, 0.0855565148647794, -0.03620067426672167, -0.006557249133752502};
static const double hmq_re_l3m2lp2_den[5] = {-6.76894063440646, 15.170831931186493, -9.406169787571082, 0.0, 1.0};
// hm_qnm.py:122
static const double hmq_re_l3m2lp3_num[7] = {1.0, -2.107852425643677, 1.1906393634562715, 0.02244848864087732, -0.09593447799423722, -0.0021343381708933025, -0.005319515989331159};
static const double hmq_re_l3m2lp3_den[7] = {1.0, -2.1078515887706324, 1.2043484690080966, 0.0, -0.08910191596778137, 0.0, -0.005471749827809503};
// hm_qnm.py:132
static const double hmq_im_l2m2lp2_num[7] = {0.0, 0.3826673013161342, -0.47531267226013896, -0.05898102880105067, 0.0724525431346487, 0.054714637311702986, 0.024544862718252784};
static const double hmq_im_l2m2lp2_den[7] = {-38.70835035062785, 69.82140084545878, -27.99036444363243, 0.0, -4.152310472191899, 0.0, 1.0};
// hm_qnm.py:142
static const double hmq_im_l2m2lp3_num[7] = {0.0, -49.7688437256778, 120.43773704442333, -82.95323455645332, 1.721453011852496, 11.540237244397877, -0.9819458637589314};
static const double hmq_im_l2m2lp3_den[7] = {2858.5790831181725, -6305.619505422591, 3825.6742092829054, 0.0, -377.7822297815406, 0.0, 1.0};
// hm_qnm.py:152
static const double hmq_im_l3m2lp2_num[7] = {0.0, -2.8704762147145533, 4.436434016918535, -1.0115343326360486, -0.08965314412106505, -0.4236810894599512, -0.041787576033810676};
static const double hmq_im_l3m2lp2_den[7] = {-171.80908957903395, 272.362882450877, -76.68544453077854, 0.0, -25.14197656531123, 0.0, 1.0};
// hm_qnm.py:162
static const double hmq_im_l3m2lp3_num[7] = {0.0, 12.45701482868677, -29.398484595717147, 18.26221675782779, 1.9308599142669403, -3.159763242921214, -0.0910871567367674};
static const double hmq_im_l3m2lp3_den[7] = {345.52914639836257, -815.4349339779621, 538.3888932415709, 0.0, -69.3840921447381, 0.0, 1.0};

// ---- hm_amp_fits_inspiral.py ----
// hm_amp_fits_inspiral.py:24
template <class R> static inline R hai_amp_21_iv1(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R sqroot = o_sqrt(eta);
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    return o_abs(o_fma(((delta * s1) * o_fma((-16.75726972301224), o_fma(876.4704270866478, eta4, o_fma((-500.4821414428368), eta3, o_fma(99.47071002831267, eta2, o_fma((-7.812073811917883), eta1, 1.1787350890261943)))), ((2.3439955698372663 * o_fma((-4768.212511142035), eta4, o_fma(2178.375177755584, eta3, o_fma((-279.6409723479635), eta2, o_fma(7.176140122833879, eta1, 0.9373952326655807))))) * s1))), sqroot, o_fma((((chidiff1 * delta) * o_fma(347.1621871204769, eta3, o_fma(32.374226994179054, eta1, (-(202.86279451816662 * eta2))))) * s1), sqroot, o_fma(((chidiff1 * delta) * o_fma(2638.3077554662905, eta3, o_fma(31.364509907424765, eta1, (-(843.6414532232126 * eta2))))), sqroot, o_fma((chidiff1 * eta5), o_fma((-134.98527058315528), s2, o_fma(987.635855365408, s1, (-3962.5020052272976))), ((delta * o_fma(220.77233521626252, eta3, o_fma((-120.58166037019478), eta2, o_fma(16.6640319856064, eta1, 19.30531354642419)))) * sqroot))))));
}

// hm_amp_fits_inspiral.py:42
template <class R> static inline R hai_amp_21_iv2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R sqroot = o_sqrt(eta);
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_abs(o_fma((delta * o_fma((chidiff1 * o_fma(748.6659668096737, eta3, o_fma(36.88578491943111, eta1, (-(321.2569602623214 * eta2))))), s1, ((chidiff1 * o_fma((-3656.192371615541), eta3, o_fma((-95.42418611585117), eta1, (1217.338674959742 * eta2)))) * s2))), sqroot, o_fma(((delta * s1) * o_fma((-10.809007068469844), o_fma((-16075.676528787526), eta5, o_fma(11501.848278358668, eta4, o_fma((-3064.7291458207815), eta3, o_fma(368.25127088896744, eta2, o_fma((-18.140470190766937), eta1, 1.1408749895922659))))), ((1.0088254664333147 * o_fma((-177907.92440833704), eta5, o_fma(130764.22485304279, eta4, o_fma((-35561.24587952242), eta3, o_fma(4257.760834055382, eta2, o_fma((-192.2461213084741), eta1, 1.2322739396680107)))))) * s1))), sqroot, o_fma((delta * o_fma((-1556.9475493549746), eta4, o_fma(1046.3187137392433, eta3, o_fma((-260.9979123967537), eta2, o_fma(30.883118528233638, eta1, 11.414934585404561))))), sqroot, o_fma((chidiff1 * eta5), o_fma(22.251142639924076, s2, o_fma(580.9465034962822, s1, (-2898.9172078672705))), ((delta * o_fma(chidiff2, o_fma((-417.5186332459383), eta3, o_fma((-18.541685007214625), eta1, (166.7427445020744 * eta2))), (chidiff1 * o_fma(2308.6520892707795, eta3, o_fma(41.61457952037761, eta1, (-(779.9151607638761 * eta2))))))) * sqroot))))));
}

// hm_amp_fits_inspiral.py:61
template <class R> static inline R hai_amp_21_iv3(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R sqroot = o_sqrt(eta);
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    return o_abs(o_fma(((delta * s1) * o_fma((-(0.28384547935698246 * o_fma((-24507.800226675925), eta4, o_fma(14158.178124971111, eta3, o_fma((-2707.9572896559484), eta2, o_fma(180.0378964169756, eta1, (-0.8870770459576875))))))), s2, o_fma((-7.1976397556450715), o_fma(1109.0598539312573, eta4, o_fma((-635.1835857158857), eta3, o_fma(120.87934060776237, eta2, o_fma((-6.587249958654174), eta1, 0.7603360145475428)))), (-((0.0811847192323969 * o_fma((-93272.17990295641), eta4, o_fma(52586.63520999897, eta3, o_fma((-9548.970156895082), eta2, o_fma(517.4039644814231, eta1, 7.951454648295709))))) * s1))))), sqroot, o_fma((((chidiff1 * delta) * o_fma(1034.4897198648962, eta3, o_fma(29.731007410186827, eta1, (-(372.09609843131386 * eta2))))) * s1), sqroot, o_fma((delta * o_fma((-3914.3390484239667), eta4, o_fma(2435.4098320959706, eta3, o_fma((-568.0027716789259), eta2, o_fma(69.01761987509603, eta1, 5.390240326328237))))), sqroot, o_fma((chidiff1 * eta5), o_fma(16.379731479465033, s2, o_fma(157.94791186394787, s1, (-2282.9983216879655))), (((chidiff1 * delta) * o_fma(1350.476411541137, eta3, o_fma(21.935833431534224, eta1, (-(460.7130131927895 * eta2))))) * sqroot))))));
}

// hm_amp_fits_inspiral.py:79
template <class R> static inline R hai_amp_33_iv1(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R sqroot = o_sqrt(eta);
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    return o_fma(((delta * s1) * o_fma((-(0.26859293613553986 * o_fma(639989.1197424605, eta5, o_fma((-462437.2690007375), eta4, o_fma(128973.72291098491, eta3, o_fma((-16664.276483466252), eta2, o_fma(888.7768796633982, eta1, (-8.029069605349488)))))))), s2, o_fma((-9.753610761811967), o_fma((-26407.40988450443), eta5, o_fma(19794.855873435758, eta4, o_fma((-5652.44754829634), eta3, o_fma(750.8933447725581, eta2, o_fma((-44.07982999150369), eta1, 1.7819678168496158))))), ((0.014210376114848208 * o_fma(7911048.294733888, eta5, o_fma((-4948175.330328345), eta4, o_fma(1145802.2059130718, eta3, o_fma((-125763.47850622259), eta2, o_fma(7264.159472864562, eta1, (-196.97328616330392))))))) * s1)))), sqroot, o_fma((((chidiff1 * delta) * o_fma(965.4670353331536, eta3, o_fma(29.221863857271703, eta1, (-(348.1658322276406 * eta2))))) * s1), sqroot, o_fma((delta * o_fma((-8113.199260593833), eta4, o_fma(4598.64854748635, eta3, o_fma((-898.470298591732), eta2, o_fma(68.14577638017842, eta1, 20.94372147101354))))), sqroot, o_fma((chidiff1 * eta5), o_fma(1.4157230717300835, s2, o_fma(26.852777193715088, s1, 155.1434307076563)), (((chidiff1 * delta) * o_fma((-141.85538063933927), eta3, o_fma(6.296698171560171, eta1, (15.81328761563562 * eta2)))) * sqroot)))));
}

// hm_amp_fits_inspiral.py:97
template <class R> static inline R hai_amp_33_iv2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R sqroot = o_sqrt(eta);
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    return o_fma(((delta * s1) * o_fma((-(0.15528289788512326 * o_fma(1451659.7366230418, eta5, o_fma((-953010.5908118751), eta4, o_fma(231384.03447562453, eta3, o_fma((-25091.14888164863), eta2, o_fma(1129.44485109116, eta1, (-12.738184090548508)))))))), s2, o_fma((-5.2580116732827085), o_fma((-36383.824902258915), eta5, o_fma(25678.53964955809, eta4, o_fma((-6879.681319382729), eta3, o_fma(861.1650630146937, eta2, o_fma((-48.20753331991333), eta1, 1.7794900975289085))))), ((0.028627002336747746 * o_fma(454655.37836367317, eta5, o_fma((-168311.2370167227), eta4, o_fma(15062.821881048358, eta3, o_fma((-2287.0465658878725), eta2, o_fma(734.7581857539398, eta1, (-50.57295946557892))))))) * s1)))), sqroot, o_fma((((chidiff1 * delta) * o_fma(327.2467193417936, eta3, o_fma(13.854211287141906, eta1, (-(135.54916401086845 * eta2))))) * s1), sqroot, o_fma((delta * o_fma((-3100.2396041211605), eta4, o_fma(1773.8883882162215, eta3, o_fma((-357.6437296928763), eta2, o_fma(35.75290806646574, eta1, 12.501397517602173))))), sqroot, o_fma((chidiff1 * eta5), o_fma((-0.16889712161410445), s2, o_fma(37.141092711336846, s1, 161.62678370819597)), (((chidiff1 * delta) * o_fma((-249.71072528701757), eta3, o_fma(3.4895829486899825, eta1, (51.07954458810889 * eta2)))) * sqroot)))));
}

// hm_amp_fits_inspiral.py:115
template <class R> static inline R hai_amp_33_iv3(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R sqroot = o_sqrt(eta);
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R eta6 = (eta1 * eta5);
    R eta7 = (eta1 * eta6);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((delta * o_fma((-521279.50781871413), eta6, o_fma(460912.76402476896, eta5, o_fma((-162498.31493332976), eta4, o_fma(29052.156454239772, eta3, o_fma((-2774.5868652930294), eta2, o_fma(148.98231189649468, eta1, 5.148353856800232))))))), sqroot, o_fma((chidiff1 * eta7), o_fma(293.89657731755483, s2, o_fma(793.5988192446642, s1, o_fma(47.94869769580204, chidiff2, 2667.3441342894776))), o_fma((delta * s1), o_fma((-1.2846335585108297), o_fma((-70.83896459998664), eta4, o_fma((-34.842921754693386), eta3, o_fma(23.26434219690476, eta2, o_fma(1.37856806162599, eta1, 0.09991079016763821)))), (-((0.03496714763391888 * o_fma((-43221.0453556626), eta4, o_fma(22714.70643022915, eta3, o_fma((-3736.1574640444287), eta2, o_fma(188.38585449575902, eta1, (-0.230558571912664)))))) * s1))), o_fma((chidiff1 * delta), o_fma((-110.10276573567405), eta3, o_fma((-0.5869777957488564), eta1, (32.65536124256588 * eta2))), (((chidiff1 * delta) * o_fma(113.77466499598913, eta3, o_fma(3.524800489907584, eta1, (-(40.26479860265549 * eta2))))) * s1)))));
}

// hm_amp_fits_inspiral.py:136
template <class R> static inline R hai_amp_32_iv1(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R sqroot = o_sqrt(eta);
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R eta6 = (eta1 * eta5);
    R s1 = s;
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((s1 * o_fma(0.6343034651912586, o_fma(30602.210950069973, eta5, o_fma((-24776.529123239357), eta4, o_fma(7616.234910399297, eta3, o_fma((-1087.6241783616488), eta2, o_fma(78.98200041834092, eta1, (-2.5844888818001737)))))), (-((0.062088720220899465 * o_fma(198366.10615196894, eta5, o_fma((-138572.93700180828), eta4, o_fma(33822.437731298516, eta3, o_fma((-3124.4712274775407), eta2, o_fma(36.01386705325694, eta1, 6.5586380356588565)))))) * s1)))), sqroot, o_fma((((chidiff1 * delta) * o_fma((-2069.5349737049073), eta4, o_fma(1160.9833883184604, eta3, o_fma(12.488240781993923, eta1, (-(209.32038774208385 * eta2)))))) * s1), sqroot, o_fma(o_fma((chidiff1 * delta), o_fma((-566.6420939162068), eta4, o_fma(278.9717709112207, eta3, o_fma((-0.739317114582042), eta1, (-(47.473246070362634 * eta2))))), (chidiff2 * o_fma(23.783684827838247, eta4, o_fma((-24.37776782232888), eta3, o_fma((-0.5873680378268906), eta1, (6.692187014925888 * eta2)))))), sqroot, (o_fma((-55436.68857586653), eta6, o_fma(51904.91007211022, eta5, o_fma((-19279.186145681153), eta4, o_fma(3585.9269057886418, eta3, o_fma((-343.3143244815765), eta2, o_fma(4.94285331708559, eta1, 3.2940434453819694)))))) * sqroot))));
}

// hm_amp_fits_inspiral.py:155
template <class R> static inline R hai_amp_32_iv2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R sqroot = o_sqrt(eta);
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R eta6 = (eta1 * eta5);
    R s1 = s;
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((s1 * o_fma(0.7718201508695763, o_fma((-970.2126139442341), eta4, o_fma(712.9101229418721, eta3, o_fma((-186.5001124789369), eta2, o_fma(26.432880113146012, eta1, (-1.3012906461000349))))), ((0.04832734931068797 * o_fma(6858.003573909322, eta4, o_fma((-2537.636347529708), eta3, o_fma(1.8360177574514709, eta2, o_fma(78.98681284391004, eta1, (-5.9999628512498315)))))) * s1))), sqroot, o_fma(o_fma(chidiff2, o_fma((-8.78587250202154), eta3, o_fma((-0.03940151060321499), eta1, (1.9034209537174116 * eta2))), ((chidiff1 * delta) * o_fma(0.36790005839460627, eta3, o_fma((-1.704299788495861), eta1, (-(4.923510922214181 * eta2)))))), sqroot, (o_fma(30950.67053883646, eta6, o_fma((-25472.73682092519), eta5, o_fma(7970.929280907627, eta4, o_fma((-1139.110227749627), eta3, o_fma(60.10368251688146, eta2, o_fma((-5.1846950040514335), eta1, 2.2911849711339123)))))) * sqroot)));
}

// hm_amp_fits_inspiral.py:174
template <class R> static inline R hai_amp_32_iv3(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R sqroot = o_sqrt(eta);
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R eta6 = (eta1 * eta5);
    R s1 = s;
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((s1 * o_fma(0.785171689871352, o_fma((-1115.008825153004), eta4, o_fma(744.7699462372949, eta3, o_fma((-167.9231114864569), eta2, o_fma(18.70856733065619, eta1, (-0.4634745514643032))))), ((0.13469300326662165 * o_fma((-472.02962341590774), eta4, o_fma(1136.8367114738041, eta3, o_fma((-483.7040402103785), eta2, o_fma(72.17373498208947, eta1, (-2.7311391326835133)))))) * s1))), sqroot, o_fma((((chidiff1 * delta) * o_fma(80.67120303906654, eta3, o_fma(1.904350804857431, eta1, (-(25.565242391371093 * eta2))))) * s1), sqroot, o_fma(o_fma(chidiff2, o_fma((-14.078156877577733), eta3, o_fma((-0.6358511175987503), eta1, (5.555088747533164 * eta2))), ((chidiff1 * delta) * o_fma(36.20685853857613, eta3, o_fma(0.23205448591711159, eta1, (-(19.46049432345157 * eta2)))))), sqroot, (o_fma(38785.53683146884, eta6, o_fma((-30849.228936891763), eta5, o_fma(8880.536804741967, eta4, o_fma((-976.9940395257111), eta3, o_fma((-17.51265776660515), eta2, o_fma(7.380126197972549, eta1, 1.1525594672495008)))))) * sqroot))));
}

// hm_amp_fits_inspiral.py:193
template <class R> static inline R hai_amp_44_iv1(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R sqroot = o_sqrt(eta);
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R eta6 = (eta1 * eta5);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((s1 * o_fma((-(0.023693246676754775 * o_fma(18834603.94974573, eta6, o_fma((-17434160.932177313), eta5, o_fma(6255666.490084672, eta4, o_fma((-1080334.0811860575), eta3, o_fma(89338.32554683842, eta2, o_fma((-2911.2154288136217), eta1, 16.437107575918503)))))))), s2, o_fma((-0.41839636169678796), o_fma((-788783.0559069642), eta6, o_fma(758333.2655404843, eta5, o_fma((-283097.33358214336), eta4, o_fma(50694.024801783446, eta3, o_fma((-4232.508174045782), eta2, o_fma(104.62892421207803, eta1, 5.143510231379954)))))), (-((0.05653522061311774 * o_fma((-1601048.9769238676), eta6, o_fma(1744031.5410044065, eta5, o_fma((-761661.9645651339), eta4, o_fma(165236.6480734229, eta3, o_fma((-17551.398321516353), eta2, o_fma(694.00652410087, eta1, 5.605483124564013))))))) * s1))))), sqroot, o_fma(o_fma((chidiff1 * delta), o_fma((-34.98399465240273), eta3, o_fma(0.5697308729057493, eta1, (8.895576813118867 * eta2))), (chidiff2 * o_fma(33.182723737396294, eta3, o_fma(1.6370346538130884, eta1, (-(14.597095790380884 * eta2)))))), sqroot, (o_fma((-923.5495700703648), eta4, o_fma(603.7453704122706, eta3, o_fma((-138.9749850448088), eta2, o_fma((-3.557926105832778), eta1, 5.2601381002242595)))) * sqroot)));
}

// hm_amp_fits_inspiral.py:213
template <class R> static inline R hai_amp_44_iv2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R sqroot = o_sqrt(eta);
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R eta6 = (eta1 * eta5);
    R s1 = s;
    R s2 = (s1 * s1);
    R s3 = (s1 * s2);
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((s1 * o_fma((-(0.027167813927224657 * o_fma(18855858.173287075, eta6, o_fma((-16270212.243584689), eta5, o_fma(5570748.509263883, eta4, o_fma((-956052.0170024392), eta3, o_fma(84126.14531363266, eta2, o_fma((-3312.7658728016568), eta1, 34.281932237450256)))))))), s3, o_fma((-(0.008830881730801855 * o_fma((-7756341.721552802), eta6, o_fma(5858455.766230802, eta5, o_fma((-1723865.744472182), eta4, o_fma(260027.4121353132, eta3, o_fma((-23328.560285901796), eta2, o_fma(1359.7883958101172, eta1, (-37.49992494976597))))))))), s2, o_fma((-0.2627203100303006), o_fma((-234487.3875219032), eta6, o_fma(227254.37897941095, eta5, o_fma((-81525.50289542897), eta4, o_fma(12546.255587592654, eta3, o_fma((-552.1725902144143), eta2, o_fma((-52.82425783851536), eta1, 6.460396349297595)))))), (-((0.008424003742397579 * o_fma((-85106628.71580401), eta6, o_fma(75396439.48937692, eta5, o_fma((-26446950.627957724), eta4, o_fma(4620165.968920881, eta3, o_fma((-408022.6805482195), eta2, o_fma(15514.571912666677, eta1, (-109.26773035716548)))))))) * s1)))))), sqroot, o_fma((((chidiff1 * delta) * o_fma(97.01712147349483, eta3, o_fma(4.749663394334708, eta1, (-(42.62996105525792 * eta2))))) * s1), sqroot, o_fma(o_fma(chidiff2, o_fma((-16.648660653220123), eta3, o_fma((-0.8318312659717388), eta1, (7.6541168007977864 * eta2))), ((chidiff1 * delta) * o_fma(5.56587823143958, eta3, o_fma(2.214478316304753, eta1, (-(7.028104574328955 * eta2)))))), sqroot, (o_fma((-876.5048298377644), eta4, o_fma(604.0067075996933, eta3, o_fma((-155.22519772642607), eta2, o_fma(6.707695566702527, eta1, 3.173191054680422)))) * sqroot))));
}

// hm_amp_fits_inspiral.py:234
template <class R> static inline R hai_amp_44_iv3(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R sqroot = o_sqrt(eta);
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R eta6 = (eta1 * eta5);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((s1 * o_fma((-(0.026407497690308382 * o_fma(128167.89151782403, eta6, o_fma(148102.4257785174, eta5, o_fma((-184787.79377504133), eta4, o_fma(66150.52694069289, eta3, o_fma((-10494.512487701073), eta2, o_fma(744.4743953122965, eta1, (-17.184685557542196))))))))), s2, o_fma((-0.15782329022461472), o_fma(307714.3954026392, eta6, o_fma((-324891.4001578353), eta5, o_fma(139319.8239886073, eta4, o_fma((-31109.307191210424), eta3, o_fma(3777.477006415653, eta2, o_fma((-229.16361598098678), eta1, 6.0309399412954345)))))), (-((0.03050157254864058 * o_fma((-11388612.30369246), eta6, o_fma(10147654.212772278, eta5, o_fma((-3558976.6538499263), eta4, o_fma(612317.1751155312, eta3, o_fma((-51213.27604422822), eta2, o_fma(1609.4251694451375, eta1, 4.232861441291087))))))) * s1))))), sqroot, o_fma((((chidiff1 * delta) * o_fma(53.35003268489743, eta3, o_fma(2.3532739003216254, eta1, (-(21.37216554136868 * eta2))))) * s1), sqroot, o_fma(o_fma((chidiff1 * delta), o_fma((-9.473290540936633), eta3, o_fma(1.4739380748149558, eta1, (0.06541707987699942 * eta2))), (chidiff2 * o_fma((-8.709159662885131), eta3, o_fma((-0.3640838331639651), eta1, (3.7369795937033756 * eta2))))), sqroot, (o_fma((-599.2709938848282), eta4, o_fma(456.78649322753824, eta3, o_fma((-139.6610487470118), eta2, o_fma(12.656614578053683, eta1, 1.7335503724888923)))) * sqroot))));
}

// ---- hm_amp_fits_intermediate.py ----
// hm_amp_fits_intermediate.py:24
template <class R> static inline R ham_amp_21_int1(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_abs(o_fma(((delta * eta1) * s1), o_fma((-(0.46006975902558517 * o_fma(731179.456702448, eta5, o_fma((-528521.6231015554), eta4, o_fma(142904.08962903373, eta3, o_fma((-17333.549622759732), eta2, o_fma(826.026625945615, eta1, (-2.1663474937625975)))))))), s2, o_fma((-8.276366844994188), o_fma((-23896.392706809565), eta5, o_fma(16770.447637953577, eta4, o_fma((-4389.799658723288), eta3, o_fma(516.7886322104276, eta2, o_fma((-24.12941323757896), eta1, 1.0677538075697492))))), (-((1.6908277400304084 * o_fma((-2085.0821513466053), eta5, o_fma(592.300353344717, eta4, o_fma((-184.13091281984674), eta3, o_fma(114.8330693231833, eta2, o_fma((-29.00026389706585), eta1, 3.4799140066657928)))))) * s1)))), o_fma((delta * eta1), o_fma((chidiff1 * o_fma(1022.8553774274128, eta3, o_fma(34.17829404207186, eta1, (-(386.34587928670015 * eta2))))), s1, ((chidiff1 * o_fma(1016.6019654342113, eta3, o_fma(56.76554600963724, eta1, (-(491.4593694689354 * eta2))))) * s2)), o_fma((chidiff1 * eta5), o_fma((-1.3098861736238572), s2, o_fma((-195.27885562092342), s1, (-1698.841763891122))), o_fma((delta * eta1), o_fma(chidiff2, o_fma(63.70715919820867, eta3, o_fma(5.159755997682368, eta1, (-(30.293198248154948 * eta2)))), (chidiff1 * o_fma(1427.5951158851076, eta3, o_fma(8.262642080222694, eta1, (-(415.88826990259116 * eta2)))))), ((delta * eta1) * o_fma(2199.614895727586, eta4, o_fma((-1614.178472020212), eta3, o_fma(447.2214642597892, eta2, o_fma((-66.46950491124205), eta1, 18.55363583212328))))))))));
}

// hm_amp_fits_intermediate.py:42
template <class R> static inline R ham_amp_21_int2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    return o_abs(o_fma(((delta * eta1) * s1), o_fma((-(0.5006664599166409 * o_fma(21974.8201355744, eta4, o_fma((-11939.621138801092), eta3, o_fma(2334.0029399672358, eta2, o_fma((-212.00728173197498), eta1, 10.289020582277626)))))), s2, o_fma((-8.097876227116853), o_fma((-1063.0692862578455), eta4, o_fma(593.8035334117192, eta3, o_fma((-116.88834714736281), eta2, o_fma(9.861355377849485, eta1, 0.6569459700232806)))), (-((1.0546375154878165 * o_fma((-7141.611333893155), eta4, o_fma(4350.442990924205, eta3, o_fma((-902.5751736558435), eta2, o_fma(65.25215540635162, eta1, 0.745557030602097))))) * s1)))), o_fma((delta * eta1), o_fma((chidiff1 * o_fma(984.6826660818275, eta3, o_fma(36.358142200869295, eta1, (-(384.2123173145321 * eta2))))), s1, ((chidiff1 * o_fma((-639.190132707115), eta3, o_fma(4.159271594881928, eta1, (105.10911749116399 * eta2)))) * s2)), o_fma((chidiff1 * eta5), o_fma((-3.3637011939285015), s2, o_fma((-102.34716990474509), s1, (-1549.6199518612063))), o_fma((delta * eta1), o_fma(12.048194546899204, eta2, o_fma((-12.783698329428516), eta1, 13.757856231617446)), (((chidiff1 * delta) * eta1) * o_fma(1333.6181181686939, eta3, o_fma(15.107530092096438, eta1, (-(416.811753638553 * eta2))))))))));
}

// hm_amp_fits_intermediate.py:59
template <class R> static inline R ham_amp_21_int3(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    return o_abs(o_fma(((delta * eta1) * s1), o_fma((-(0.4308267743835775 * o_fma(29480.158198408903, eta4, o_fma((-17712.883990278668), eta3, o_fma(3662.099161055873, eta2, o_fma((-302.9708323417439), eta1, 9.970654092010587)))))), s2, o_fma((-8.665646058245033), o_fma((-1001.2370065269745), eta4, o_fma(576.7172598056907, eta3, o_fma((-111.70764910503321), eta2, o_fma(8.293609541933655, eta1, 0.7862132291286934)))), (-((0.9459820574514348 * o_fma((-7518.309520232795), eta4, o_fma(4331.56002883546, eta3, o_fma((-817.7854010574645), eta2, o_fma(48.94077040282239, eta1, 1.309016452198605))))) * s1)))), o_fma((chidiff1 * eta5), o_fma(3.0053086651515843, s2, o_fma((-16.639896279650678), s1, (-1522.8543551416456))), o_fma((((chidiff1 * delta) * eta1) * o_fma(343.79587860999874, eta3, o_fma(11.877985158418596, eta1, (-(131.04937626836355 * eta2))))), s1, o_fma((delta * eta1), o_fma((-127.62161159798488), eta3, o_fma(76.14884211156267, eta2, o_fma((-21.755549987331054), eta1, 13.318990196097973))), (((chidiff1 * delta) * eta1) * o_fma(1366.2408490833282, eta3, o_fma(17.704321326939414, eta1, (-(434.4390350012534 * eta2))))))))));
}

// hm_amp_fits_intermediate.py:76
template <class R> static inline R ham_amp_21_int4(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    return o_abs(o_fma(((delta * eta1) * s1), o_fma((-(0.21540359093306097 * o_fma((-9835.416414460055), eta4, o_fma(1889.9613367802453, eta3, o_fma(385.8345793110262, eta2, o_fma((-109.06544597367301), eta1, 9.42805409480658)))))), s2, o_fma((-9.499741513411829), o_fma((-248.5050377498615), eta4, o_fma(166.04254881175257, eta3, o_fma((-33.651192908287236), eta2, o_fma(2.400945118514037, eta1, 0.912120958549489)))), (-((0.7850652143322492 * o_fma((-9077.037917192794), eta4, o_fma(5381.481380750608, eta3, o_fma((-1032.1319480683567), eta2, o_fma(60.81773903539479, eta1, 1.534131218043425))))) * s1)))), o_fma((chidiff1 * eta5), o_fma((-2.2456928156392197), s2, o_fma(74.44462583487345, s1, (-1528.956015503355))), o_fma((((chidiff1 * delta) * eta1) * o_fma(280.6003655502327, eta3, o_fma(12.925363020014743, eta1, (-(126.87194512915104 * eta2))))), s1, o_fma((delta * eta1), o_fma((-139.25546924151664), eta3, o_fma(83.20619262213437, eta2, o_fma((-22.831152256559523), eta1, 13.094382343446163))), (((chidiff1 * delta) * eta1) * o_fma(1430.3698681181, eta3, o_fma(20.120192352555357, eta1, (-(458.2592421214168 * eta2))))))))));
}

// hm_amp_fits_intermediate.py:93
template <class R> static inline R ham_amp_33_int1(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    return o_fma((chidiff1 * eta5), o_fma(7.650565415855383, s2, o_fma(53.36849263931051, s1, 199.6863414922219)), o_fma(((delta * eta1) * s1), o_fma((-0.5383304368673182), o_fma(11959.83533107835, eta5, o_fma((-9675.194644814854), eta4, o_fma(3507.3655567272644, eta3, o_fma((-843.7897535238325), eta2, o_fma(129.36947401891433, eta1, (-7.456619067234563)))))), (-((0.28042799223829407 * o_fma(166056.27237509796, eta5, o_fma((-119209.70783201039), eta4, o_fma(32634.43965039936, eta3, o_fma((-4241.537539226717), eta2, o_fma(266.69059813274475, eta1, (-6.212827413930676))))))) * s1))), o_fma((((chidiff1 * delta) * eta1) * o_fma(99.91399033058927, eta3, o_fma(5.575274516197629, eta1, (-(44.592719238427094 * eta2))))), s1, o_fma(((chidiff1 * delta) * eta1), o_fma((-148.3162618111991), eta3, o_fma((-0.3516244197696068), eta1, (40.425151307421416 * eta2))), ((delta * eta1) * o_fma(3996.417635728702, eta4, o_fma((-3047.513201789169), eta3, o_fma(914.5350366065115, eta2, o_fma((-146.29035440932105), eta1, 26.998512565991778)))))))));
}

// hm_amp_fits_intermediate.py:110
template <class R> static inline R ham_amp_33_int2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    return o_fma(((delta * eta1) * s1), o_fma((0.30983324991828787 * o_fma((-504089.07690692553), eta5, o_fma(376409.2335857397, eta4, o_fma((-104863.50453518033), eta3, o_fma(13278.047199998868, eta2, o_fma((-719.9854052004307), eta1, 11.302200127272357)))))), s2, o_fma(0.3104472390387834, o_fma(205525.13910847943, eta5, o_fma((-154255.3408672655), eta4, o_fma(42566.59565666228, eta3, o_fma((-4964.199967099143), eta2, o_fma(169.85423386969634, eta1, 6.073591341439855))))), ((0.2295327944679772 * o_fma(79210.33893514868, eta5, o_fma((-41567.07396803811), eta4, o_fma(2404.4151687877525, eta3, o_fma(1876.408148917458, eta2, o_fma((-354.7914372697625), eta1, 19.236275867648594)))))) * s1))), o_fma((chidiff1 * eta5), o_fma((-10.05114600643587), s2, o_fma(51.2153901931183, s1, 234.13476431269862)), o_fma((((chidiff1 * delta) * eta1) * o_fma(60.38649425562178, eta3, o_fma((-0.3265505633310564), eta1, (-(9.861644053348053 * eta2))))), s1, o_fma((delta * eta1), o_fma(50.726220750178435, eta2, o_fma((-28.970875603981295), eta1, 17.42562079069636)), (((chidiff1 * delta) * eta1) * o_fma((-273.1170921735085), eta3, o_fma((-7.861956897615623), eta1, (93.45476935080045 * eta2))))))));
}

// hm_amp_fits_intermediate.py:127
template <class R> static inline R ham_amp_33_int3(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    return o_fma(((delta * eta1) * s1), o_fma((0.3695850566805098 * o_fma((-534437.8832860565), eta5, o_fma(360950.6686625352, eta4, o_fma((-91096.64982858274), eta3, o_fma(10426.348944657859, eta2, o_fma((-498.95526910874986), eta1, 6.077183107132255)))))), s2, o_fma(0.4590550434774332, o_fma(282064.174913521, eta5, o_fma((-195187.55422847517), eta4, o_fma(50003.94075934942, eta3, o_fma((-5545.0086137386525), eta2, o_fma(207.74562213604057, eta1, 3.0594364612798635))))), ((0.657748992123043 * o_fma((-20879.999628681435), eta5, o_fma(19714.675715229736, eta4, o_fma((-6999.7659193505915), eta3, o_fma(1276.6209573025596, eta2, o_fma((-124.06189543062042), eta1, 5.57939137343977)))))) * s1))), o_fma((chidiff1 * eta5), o_fma(10.457035444964653, s2, o_fma(44.408527093031566, s1, 263.5151703818307)), o_fma((((chidiff1 * delta) * eta1) * o_fma((-73.62033985436068), eta3, o_fma((-7.048898856045782), eta1, (49.941617405768135 * eta2)))), s1, o_fma((delta * eta1), o_fma(16.79500349318081, eta2, o_fma((-12.799844096694798), eta1, 14.555522136327964)), (((chidiff1 * delta) * eta1) * o_fma((-562.0936797781519), eta3, o_fma((-16.292654447108134), eta1, (190.3516012682791 * eta2))))))));
}

// hm_amp_fits_intermediate.py:144
template <class R> static inline R ham_amp_33_int4(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma(((delta * eta1) * s1), o_fma((0.2181399561677432 * o_fma(277256.38289831823, eta5, o_fma((-216137.90827404748), eta4, o_fma(59288.71069769708, eta3, o_fma((-6857.936471898544), eta2, o_fma(303.9394443302189, eta1, (-1.6754158383043574))))))), s2, o_fma(0.6893003654021495, o_fma((-27679.226582207506), eta5, o_fma(19514.812242035154, eta4, o_fma((-4797.429130246123), eta3, o_fma(565.3767256471909, eta2, o_fma((-44.83989278653052), eta1, 3.1014226377197027))))), ((0.7068016563068026 * o_fma((-65658.74746265226), eta5, o_fma(48603.96661003743, eta4, o_fma((-13485.270489656365), eta3, o_fma(1788.1730303291356, eta2, o_fma((-118.51094098279343), eta1, 4.071212304920691)))))) * s1))), o_fma((chidiff1 * eta5), o_fma((-29.471275170013012), s2, o_fma(7.635582066682054, s1, o_fma((-21.102665129433042), chidiff2, 287.90444670305715))), o_fma((((chidiff1 * delta) * eta1) * o_fma((-1485.5134187612182), eta3, o_fma((-46.85669289852532), eta1, (522.3965959942979 * eta2)))), s1, o_fma((delta * eta1), o_fma(17.098576301150125, eta2, o_fma((-7.449975618083432), eta1, 13.312095699772305)), ((delta * eta1) * o_fma(chidiff1, o_fma((-1103.1917047361735), eta3, o_fma((-31.171150896110156), eta1, (371.1389274783572 * eta2))), (chidiff2 * o_fma(1164.9282236341376, eta3, o_fma(32.78644599730888, eta1, (-(395.15713118955387 * eta2)))))))))));
}

// hm_amp_fits_intermediate.py:162
template <class R> static inline R ham_amp_32_int1(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R sqroot = o_sqrt(eta);
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((o_fma(106683.98302194536, eta4, o_fma((-99873.4707358776), eta3, o_fma(480.0434256230109, eta1, (25346.341240810478 * eta2)))) * sqroot), (1.0 / o_fma(10083.297670051445, eta2, o_fma(1082.6574834474493, eta1, 1.0))), o_fma((s1 * o_fma(0.7067243321652764, o_fma((-1929.9084724384807), eta5, o_fma(175.4194342269804, eta4, o_fma(325.71940024255775, eta3, o_fma((-85.46581740280585), eta2, o_fma(9.608999847549535, eta1, 0.12885110296881636))))), ((0.1540566313813899 * o_fma(43309.69769878964, eta5, o_fma((-28804.241518798244), eta4, o_fma(7184.647314370326, eta3, o_fma((-827.591235943271), eta2, o_fma(45.55785402900492, eta1, (-0.3261041495083288))))))) * s1))), sqroot, o_fma(o_fma(chidiff2, o_fma((-8.68296921440857), eta3, o_fma((-0.2341404256829785), eta1, (2.606326837996192 * eta2))), ((chidiff1 * delta) * o_fma(73.40268975811729, eta3, o_fma(0.5454562486736877, eta1, (-(25.19759222940851 * eta2)))))), sqroot, ((((chidiff1 * delta) * o_fma(32.22238925527844, eta3, o_fma(0.4422257616009941, eta1, (-(8.490112284851655 * eta2))))) * s1) * sqroot))));
}

// hm_amp_fits_intermediate.py:180
template <class R> static inline R ham_amp_32_int2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R eta6 = (eta1 * eta5);
    R s1 = s;
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((eta1 * s1), o_fma(1.5098291294292217, o_fma((-1633.3295758142904), eta4, o_fma(856.9917885742416, eta3, o_fma((-143.66427232396376), eta2, o_fma(9.848766999273414, eta1, 0.4844667556328104)))), ((0.32413108737204144 * o_fma(6660.304740652403, eta4, o_fma((-3811.5047139343505), eta3, o_fma(761.6103793011912, eta2, o_fma((-62.37317183581803), eta1, 2.835358206961064))))) * s1)), o_fma((((chidiff1 * delta) * eta1) * o_fma(68.72798751937934, eta3, o_fma(1.2198776533959694, eta1, (-(26.816651899746475 * eta2))))), s1, o_fma(eta1, o_fma(chidiff2, o_fma((-128.88334273588077), eta3, o_fma((-4.175680729484314), eta1, (47.54281549129226 * eta2))), ((chidiff1 * delta) * o_fma(208.07105580635888, eta3, o_fma((-0.18274358639599947), eta1, (-(71.01128541687838 * eta2)))))), (eta1 * o_fma(77839.74180254337, eta6, o_fma((-69430.10468748478), eta5, o_fma(24956.9592553473, eta4, o_fma((-4544.076411013166), eta3, o_fma(456.2188780552874, eta2, o_fma((-38.57900689641654), eta1, 4.760999387359598))))))))));
}

// hm_amp_fits_intermediate.py:198
template <class R> static inline R ham_amp_32_int3(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R eta6 = (eta1 * eta5);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma(s1, o_fma((-(0.03985156529181297 * o_fma(162398.81483375572, eta6, o_fma((-46454.57664737511), eta5, o_fma((-7859.4784979916085), eta4, o_fma((-304.92981902871617), eta2, (3614.518459296278 * eta3))))))), s2, o_fma(0.3650871458400108, o_fma(16001.318492586077, eta6, o_fma((-18580.859291282686), eta5, o_fma(6734.058809060536, eta4, o_fma(71.57390929624825, eta2, (-(994.5272351916166 * eta3)))))), ((0.0960146077440495 * o_fma(451546.88775684836, eta6, o_fma((-318877.43061174755), eta5, o_fma(83403.5743434538, eta4, o_fma(451.74917589707513, eta2, (-(9719.470997418284 * eta3))))))) * s1))), o_fma(((chidiff1 * delta) * o_fma((-26.46699529970884), eta4, o_fma((-0.2566958540737833), eta2, (14.488550203412675 * eta3)))), s1, o_fma(403.5748987560612, eta5, o_fma((chidiff1 * delta), o_fma(387.00162600306226, eta4, o_fma(1.924723094787216, eta2, (-(90.92716917757797 * eta3)))), o_fma((-58.408542930248046), eta4, o_fma(chidiff2, o_fma((-203.65975422378446), eta4, o_fma((-7.065118823041031), eta2, (77.97950589523865 * eta3))), o_fma(1.7262466525848588, eta3, o_fma(3.881450518842405, eta1, (-(12.580316392558837 * eta2))))))))));
}

// hm_amp_fits_intermediate.py:217
template <class R> static inline R ham_amp_32_int4(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R s1 = s;
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((eta1 * s1), o_fma(1.242750265695504, o_fma((-3797.3561904195085), eta4, o_fma(1858.1518762309654, eta3, o_fma((-303.5223688400309), eta2, o_fma(25.168480118950065, eta1, (-0.522172424518215))))), ((0.2927045241764365 * o_fma(6097.887891566872, eta4, o_fma((-3131.5783196211587), eta3, o_fma(471.64047356915603, eta2, o_fma((-15.488754837330958), eta1, 0.5056957789079993))))) * s1)), o_fma((((chidiff1 * delta) * eta1) * o_fma((-144.21651114700492), eta3, o_fma((-3.582040878719185), eta1, (57.75888914133383 * eta2)))), s1, o_fma(eta1, o_fma(chidiff2, o_fma((-236.2438921965621), eta3, o_fma((-8.572797326909152), eta1, (92.95723645687826 * eta2))), ((chidiff1 * delta) * o_fma(645.2760206304703, eta3, o_fma(6.674358856924571, eta1, (-(171.4826985994883 * eta2)))))), (eta1 * o_fma(692.926425981414, eta4, o_fma((-143.41033155133266), eta3, o_fma(25.254017911686333, eta2, o_fma((-16.57299637423352), eta1, 3.921660532875504))))))));
}

// hm_amp_fits_intermediate.py:233
template <class R> static inline R ham_amp_44_int1(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((eta1 * s1), o_fma((-(0.07704777584463054 * o_fma((-858.9168637051405), eta3, o_fma(344.9177692251726, eta2, o_fma((-50.35070009227704), eta1, 4.581767671445529))))), s2, o_fma((-0.062333275821238224), o_fma((-1013.1110639471394), eta3, o_fma(64.92115530780129, eta2, o_fma(137.4395894877131, eta1, (-21.630297087123807)))), (-((0.11014697070998722 * o_fma((-1828.2527520190122), eta3, o_fma(831.6073263887092, eta2, o_fma((-108.6912882442823), eta1, 4.149721483857751)))) * s1)))), o_fma(eta1, o_fma((chidiff1 * delta), o_fma(16.879953490422782, eta3, o_fma(1.5378890240544967, eta1, (-(3.4499418893734903 * eta2)))), (chidiff2 * o_fma(23.259283336239545, eta3, o_fma(1.720226708214248, eta1, (-(11.87925165364241 * eta2)))))), (eta1 * o_fma(1575.602443847111, eta4, o_fma((-1111.9864921907126), eta3, o_fma(324.1998823562892, eta2, o_fma((-64.95499142822892), eta1, 8.790173464969538)))))));
}

// hm_amp_fits_intermediate.py:250
template <class R> static inline R ham_amp_44_int2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R eta6 = (eta1 * eta5);
    R s1 = s;
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((eta1 * s1), o_fma((-0.07801583359561987), o_fma((-702172.2934567826), eta6, o_fma(760889.6919776166, eta5, o_fma((-350025.2112501252), eta4, o_fma(83447.0461509547, eta3, o_fma((-10671.072606753183), eta2, o_fma(752.1603553640072, eta1, (-28.29972282146242))))))), ((0.013159545629626014 * o_fma(268756.85659532435, eta6, o_fma((-1150518.6449682908), eta5, o_fma(1014187.7915334814, eta4, o_fma((-344254.9973814295), eta3, o_fma(52391.684517955284, eta2, o_fma((-3557.5003799977294), eta1, 91.1469833190294))))))) * s1)), o_fma(eta1, o_fma((chidiff1 * delta), o_fma(44.78225529547671, eta3, o_fma(2.3123974306694057, eta1, (-(12.237594841284904 * eta2)))), (chidiff2 * o_fma(61.05270871360041, eta3, o_fma(2.9282931698944292, eta1, (-(25.624210264341933 * eta2)))))), (eta1 * o_fma(1478.050456337336, eta4, o_fma((-920.358408667518), eta3, o_fma(236.76146303619544, eta2, o_fma((-46.81443520117986), eta1, 6.98072197826729)))))));
}

// hm_amp_fits_intermediate.py:268
template <class R> static inline R ham_amp_44_int3(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((eta1 * s1), o_fma((-(0.060477799540741804 * o_fma(426.0861915646515, eta3, o_fma((-569.4799973930861), eta2, o_fma(175.6116682068523, eta1, (-14.578189130145661)))))), s2, o_fma((-0.23316184683282615), o_fma((-1115.4147472977265), eta3, o_fma(409.3387618483284, eta2, o_fma((-23.606399143993716), eta1, (-1.7279023138971559)))), (-((0.09653777612560172 * o_fma((-1815.7529908827103), eta3, o_fma(541.1248219096527, eta2, o_fma((-2.1852511802701264), eta1, (-5.310643306559746))))) * s1)))), o_fma(eta1, o_fma((chidiff1 * delta), o_fma((-43.675503209991184), eta3, o_fma((-0.8765502142143329), eta1, (22.806632458441996 * eta2))), (chidiff2 * o_fma(16.18571810759235, eta3, o_fma(0.48698617426180074, eta1, (-(4.302527065360426 * eta2)))))), (eta1 * o_fma(2379.538739132234, eta4, o_fma((-1285.7635006711453), eta3, o_fma(269.44092930942793, eta2, o_fma((-44.10631039734796), eta1, 6.379772583015967)))))));
}

// hm_amp_fits_intermediate.py:285
template <class R> static inline R ham_amp_44_int4(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((eta1 * s1), o_fma((-(0.06565232123453196 * o_fma(1044228.2862506858, eta5, o_fma((-845341.6001856657), eta4, o_fma(253694.6032002248, eta3, o_fma((-33951.35186039629), eta2, o_fma(1869.0168486099005, eta1, (-26.15227471380236)))))))), s2, o_fma((-0.36068516166901304), o_fma((-91269.3745059139), eta5, o_fma(60287.45399959349, eta4, o_fma((-14925.514654896673), eta3, o_fma(1618.4222330016048, eta2, o_fma((-47.56175350408845), eta1, (-2.120354236840677)))))), (-((0.09635801207669747 * o_fma((-67024.98285179552), eta5, o_fma((-4102.218189945819), eta4, o_fma(16655.87939259747, eta3, o_fma((-4176.398139238679), eta2, o_fma(371.7551657959369, eta1, (-11.824692837267394))))))) * s1)))), o_fma(eta1, o_fma((chidiff1 * delta), o_fma((-112.2714974622516), eta3, o_fma((-2.461738962276138), eta1, (45.3240543970684 * eta2))), (chidiff2 * o_fma(28.44633544874233, eta3, o_fma(0.9158352037567031, eta1, (-(8.724582331021695 * eta2)))))), (eta1 * o_fma(4067.1685640401033, eta4, o_fma((-2002.2013283876834), eta3, o_fma(350.97192927929433, eta2, o_fma((-45.42463610529546), eta1, 6.098676337298138)))))));
}

// ---- hm_amp_fits_ringdown.py ----
// hm_amp_fits_ringdown.py:19
template <class R> static inline R har_amp_21_rdcp1(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R eta6 = (eta1 * eta5);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    return o_abs(o_fma((chidiff1 * eta5), o_fma(10.13490934572329, s2, o_fma(175.59081294852444, s1, (-1512.439342647443))), o_fma(((delta * eta1) * s1), o_fma((-10.090867481062709), o_fma((-1992.9074348833092), eta6, o_fma(3950.7894052628735, eta5, o_fma((-2688.211889175019), eta4, o_fma(801.3948756800821, eta3, o_fma((-107.22993216128548), eta2, o_fma(5.388149112485179, eta1, 0.9580746052260011)))))), (-((0.42972412296628143 * o_fma(658866.0983367155, eta6, o_fma((-479649.75188253267), eta5, o_fma(107980.65459735804, eta4, o_fma((-3176.4950303461164), eta3, o_fma((-1616.9974609915555), eta2, o_fma(139.73149069609775, eta1, 1.9193131231064235))))))) * s1))), o_fma((((chidiff1 * delta) * eta1) * o_fma(591.1097623278739, eta3, o_fma(21.402708785047853, eta1, (-(232.07306353130417 * eta2))))), s1, o_fma((delta * eta1), o_fma((-175.16681482428694), eta3, o_fma(92.6090002736012, eta2, o_fma((-23.5291063016996), eta1, 12.880905080761432))), (((chidiff1 * delta) * eta1) * o_fma(2255.040486907459, eta3, o_fma(26.89427230731867, eta1, (-(710.8871223808559 * eta2))))))))));
}

// hm_amp_fits_ringdown.py:37
template <class R> static inline R har_amp_21_rdcp2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    return o_abs(o_fma(((delta * eta1) * s1), o_fma((0.40226131643223145 * o_fma((-29331.524552164105), eta4, o_fma(16087.433824017446, eta3, o_fma((-2885.5110686004946), eta2, o_fma(190.42214703482531, eta1, (-3.834742668014861)))))), s2, o_fma((-8.236485576091717), o_fma((-997.6961311502365), eta4, o_fma(519.0146702141192, eta3, o_fma((-87.05136337926156), eta2, o_fma(5.1501231412520285, eta1, 0.8917610178208336)))), ((0.2836840678615208 * o_fma(2330.3534917059906, eta4, o_fma((-1882.2040277496196), eta3, o_fma(586.7942442434971, eta2, o_fma((-57.65586769647737), eta1, (-0.19281297100324718)))))) * s1))), o_fma((chidiff1 * eta5), o_fma(21.159179960295855, s2, o_fma(245.73200219767514, s1, (-1116.4039232324135))), o_fma((((chidiff1 * delta) * eta1) * o_fma((-20.274691328125606), eta3, o_fma(2.7866868976718226, eta1, (-(0.4210629980868266 * eta2))))), s1, o_fma((delta * o_fma((-7.5304766811877455), eta1, 9.112452928978168)), eta1, (((chidiff1 * delta) * eta1) * o_fma(1618.0818430353293, eta3, o_fma(16.236533863306132, eta1, (-(500.11964987628926 * eta2))))))))));
}

// hm_amp_fits_ringdown.py:54
template <class R> static inline R har_amp_21_rdcp3(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    return o_abs(o_fma(((delta * eta1) * s1), o_fma((0.1930861073906724 * o_fma(485.67801863289293, eta3, o_fma((-78.07256429516346), eta2, o_fma((-19.11543562444476), eta1, 3.7735099269174106)))), s2, o_fma((-2.7294297839371824), o_fma((-87.26454353763077), eta3, o_fma(36.120093043420326, eta2, o_fma((-4.384077347340523), eta1, 1.148166706456899))), ((0.23949142867803436 * o_fma(862.3123076782916, eta3, o_fma((-307.3404155231787), eta2, o_fma(33.33372867559165, eta1, (-0.6931516433988293))))) * s1))), o_fma((chidiff1 * eta5), o_fma((-5.353399481074393), s2, o_fma(92.91301644484749, s1, (-334.0969956136684))), o_fma((((chidiff1 * delta) * eta1) * o_fma((-0.7347258030219584), eta3, o_fma(1.9701247529688362, eta1, (-(2.8616711550845575 * eta2))))), s1, o_fma((delta * o_fma((-3.038523690239521), eta1, 2.920930733198033)), eta1, (((chidiff1 * delta) * eta1) * o_fma(544.1978232314333, eta3, o_fma(6.3472251472354975, eta1, (-(171.23657247338042 * eta2))))))))));
}

// hm_amp_fits_ringdown.py:71
template <class R> static inline R har_amp_33_rdcp1(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma(((((13.499999999999998 * delta) * eta1) * s1) * o_fma((-0.13654149379906394), o_fma((-3591.1757995710486), eta5, o_fma((-6167.08766058184), eta4, o_fma(4142.974510926698, eta3, o_fma((-742.1357702210267), eta2, o_fma(29.023992126142304, eta1, 2.719687834084113))))), (-((0.06248535354306988 * o_fma((-10056.98422430965), eta5, o_fma(6601.918552659412, eta4, o_fma((-1907.008984765889), eta3, o_fma(444.79350113344543, eta2, o_fma((-78.23231700361792), eta1, 6.697567446351289)))))) * s1)))), (1.0 / ((-3.9329308614837704) + s1)), o_fma((chidiff1 * eta5), o_fma((-15.81624534976308), s2, o_fma((-21.836905132600755), s1, o_fma((-34.5074027042393), chidiff2, 257.33227387984863))), o_fma((((chidiff1 * delta) * eta1) * o_fma((-1653.496174539196), eta3, o_fma((-52.85961155799673), eta1, (577.6347407795782 * eta2)))), s1, o_fma((delta * eta1), o_fma(22.780673360839497, eta2, o_fma((-4.436329538596615), eta1, 12.439702602599235)), ((delta * eta1) * o_fma(chidiff1, o_fma((-1524.2981907688634), eta3, o_fma((-41.04442169938298), eta1, (502.9246970179746 * eta2))), (chidiff2 * o_fma(1020.6734178547847, eta3, o_fma(32.23960072974939, eta1, (-(365.1526474476759 * eta2)))))))))));
}

// hm_amp_fits_ringdown.py:89
template <class R> static inline R har_amp_33_rdcp2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((chidiff1 * eta5), o_fma(41.321480132899524, s2, o_fma((-1.873799747678187), s1, o_fma((-29.10412708633528), chidiff2, 159.72371180117415))), o_fma(((delta * eta1) * s1), o_fma(0.1206817303851239, o_fma((-265581.70142471837), eta5, o_fma(163644.2192178668, eta4, o_fma((-35378.156133055556), eta3, o_fma(3188.189172446398, eta2, o_fma((-144.08062755162752), eta1, 8.667503604073314))))), ((0.08028332044013944 * o_fma((-148782.33687815256), eta5, o_fma(121293.97832549023, eta4, o_fma((-35625.58409457366), eta3, o_fma(4777.45310151897, eta2, o_fma((-322.95832000179297), eta1, 12.632478544060636)))))) * s1)), o_fma((((chidiff1 * delta) * eta1) * o_fma((-636.4802385540647), eta3, o_fma((-17.18006888428382), eta1, (190.73514518113845 * eta2)))), s1, o_fma((delta * eta1), o_fma(4.543696144846763, eta1, 8.425057692276933), (((chidiff1 * delta) * eta1) * o_fma((-1293.422289802462), eta3, o_fma((-32.18860840414171), eta1, (412.07321398189293 * eta2))))))));
}

// hm_amp_fits_ringdown.py:107
template <class R> static inline R har_amp_33_rdcp3(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma(((((13.499999999999998 * delta) * eta1) * s1) * o_fma(0.0011484326782460882, o_fma((-517034.11171277676), eta5, o_fma(316864.0589150127, eta4, o_fma((-67368.27397765424), eta3, o_fma(5709.239330076732, eta2, o_fma((-172.58816646768219), eta1, 4.1815722950796035))))), (-((0.009496797093329243 * o_fma((-140089.40725211444), eta5, o_fma(94309.57566199072, eta4, o_fma((-23464.64953722729), eta3, o_fma(2628.6024206791726, eta2, o_fma((-118.35865186626413), eta1, 0.9233282181397624)))))) * s1)))), (1.0 / (o_fma((-0.41099904730526465), s1, 0.09549360183532198) + s2)), o_fma((chidiff1 * eta5), o_fma(67.23189816732847, s2, o_fma(35.09131921815571, s1, o_fma((-3.071422453072518), chidiff2, 32.07819766300362))), o_fma((((chidiff1 * delta) * eta1) * o_fma((-473.30988537734174), eta3, o_fma((-10.27418232676514), eta1, (136.5150165348149 * eta2)))), s1, o_fma((delta * eta1), o_fma(2.321696430921996, eta1, 2.485784720088995), ((delta * eta1) * o_fma(chidiff1, o_fma((-496.1564538739011), eta3, o_fma((-10.454376404653859), eta1, (147.10344302665484 * eta2))), (chidiff2 * o_fma((-197.51205149250532), eta3, o_fma((-5.9236399792925996), eta1, (65.86115501723127 * eta2))))))))));
}

// hm_amp_fits_ringdown.py:125
template <class R> static inline R har_amp_44_rdcp1(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R eta6 = (eta1 * eta5);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((eta1 * s1), o_fma((0.005856387555754387 * o_fma((-3236985.8387064277), eta5, o_fma(3054333.578686424, eta4, o_fma((-1046126.8207440731), eta3, o_fma(156431.3764198244, eta2, o_fma((-9560.410655118145), eta1, 186.39698091707513)))))), s2, o_fma((-0.4360849737360132), o_fma((-111910.71267453219), eta5, o_fma(71236.86532610047, eta4, o_fma((-16718.425586396803), eta3, o_fma(1729.1839588870455, eta2, o_fma((-58.70494649755802), eta1, (-0.9543114627170375)))))), (-((0.024861802943501172 * o_fma((-581776.5303770923), eta5, o_fma(168937.32229060197, eta4, o_fma(35332.328181283, eta3, o_fma((-15866.093368857853), eta2, o_fma(1585.462602954658, eta1, (-52.25045490410733))))))) * s1)))), o_fma(eta1, o_fma((chidiff1 * delta), o_fma((-297.2592736781142), eta3, o_fma((-8.51952446214978), eta1, (117.76530248141987 * eta2))), (chidiff2 * o_fma((-8.017569928870929), eta3, o_fma((-0.2750098647982238), eta1, (4.456900599347149 * eta2))))), (eta1 * o_fma(158292.58792484726, eta6, o_fma((-98374.18361532023), eta5, o_fma(25108.811524802128, eta4, o_fma((-3514.3385364216438), eta3, o_fma(287.9418482197136, eta2, o_fma((-33.67252878543393), eta1, 5.635069974807398)))))))));
}

// hm_amp_fits_ringdown.py:144
template <class R> static inline R har_amp_44_rdcp2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((eta1 * s1), o_fma((0.019740635678180102 * o_fma(910729.0219043035, eta5, o_fma((-696891.1349708183), eta4, o_fma(201981.0743810629, eta3, o_fma((-27212.67022616794), eta2, o_fma(1535.781942095697, eta1, (-7.046295936301379))))))), s2, o_fma((-0.23311779185707152), o_fma((-127148.27131388368), eta5, o_fma(73397.72190288734, eta4, o_fma((-14672.23221502075), eta3, o_fma(1163.9107546486134, eta2, o_fma((-20.12558747513885), eta1, (-1.0795711755430002)))))), ((0.025805905356653 * o_fma(1323862.4538095295, eta5, o_fma((-819148.9390278616), eta4, o_fma(174164.91607515427, eta3, o_fma((-14580.02701600596), eta2, o_fma(350.93274421955806, eta1, 11.929946153728276)))))) * s1))), o_fma(eta1, o_fma((chidiff1 * delta), o_fma((-123.94152825700999), eta3, o_fma((-2.861653255976984), eta1, (50.50227103211222 * eta2))), (chidiff2 * o_fma(72.40230240887851, eta3, o_fma(2.9415751419018865, eta1, (-(28.79779545444817 * eta2)))))), (eta1 * o_fma(49148.57065911245, eta5, o_fma((-30595.4993786313), eta4, o_fma(7168.843978909433, eta3, o_fma((-792.0167314124681), eta2, o_fma(25.15310593958783, eta1, 3.2461722686239307))))))));
}

// hm_amp_fits_ringdown.py:162
template <class R> static inline R har_amp_44_rdcp3(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma((eta1 * s1), o_fma((-0.048465588779596405), o_fma((-133340.93723954144), eta5, o_fma(76620.18243090731, eta4, o_fma((-16522.343895064576), eta3, o_fma(1744.329802302927, eta2, o_fma((-81.33156665674845), eta1, (-0.34041762314288154)))))), ((0.024804027856323612 * o_fma(584744.0406581408, eta5, o_fma((-422282.0368440555), eta4, o_fma(112832.04975245205, eta3, o_fma((-13644.988225595187), eta2, o_fma(711.8727878341302, eta1, (-8.666095805675418))))))) * s1)), o_fma(eta1, o_fma((chidiff1 * delta), o_fma(66.58241012629095, eta3, o_fma(2.4286414692113816, eta1, (-(23.213332913737403 * eta2)))), (chidiff2 * o_fma(78.49621016381445, eta3, o_fma(3.085167288859442, eta1, (-(31.60440418701438 * eta2)))))), (eta1 * o_fma(18536.937955014455, eta5, o_fma((-12028.786386004691), eta4, o_fma(2932.3415281149432, eta3, o_fma((-337.70598252897696), eta2, o_fma(13.695204704208976, eta1, 0.861883217178703))))))));
}

// hm_amp_fits_ringdown.py:179
template <class R> static inline R har_amp_32_alambda(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R eta6 = (eta1 * eta5);
    R s1 = s;
    R s2 = (s1 * s1);
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma(o_fma(14.163264397061505, eta5, o_fma((-15.707327897569396), eta4, o_fma(6.5972460654253515, eta3, o_fma(0.08408469319155859, eta1, (-(1.223794846617597 * eta2)))))), (1.0 / o_fma(18.93655612952139, eta2, o_fma((-8.612447115134758), eta1, 1.0))), o_fma(s1, o_fma((-(0.00043088431510840695 * o_fma((-6094089.675611567), eta6, o_fma(3907267.4134789007, eta5, o_fma((-870035.1394696251), eta4, o_fma(76517.21042363928, eta3, o_fma((-30.014415072587354), eta1, (-(1900.5495690280086 * eta2))))))))), s2, o_fma((-0.0013330716557843666), o_fma((-150504.2927996263), eta6, o_fma(252228.94931931415, eta5, o_fma((-102969.88022112886), eta4, o_fma(16509.69605686402, eta3, o_fma(56.35538385647113, eta1, (-(1218.1550992423377 * eta2))))))), ((0.0010126460331462495 * o_fma((-597155.7619486618), eta6, o_fma(295469.28617550555, eta5, o_fma((-36980.079328277505), eta4, o_fma((-1304.9210590188136), eta3, o_fma((-33.87083889060834), eta1, (502.6221651850776 * eta2))))))) * s1))), o_fma(((chidiff1 * delta) * o_fma(117.66195692191727, eta5, o_fma(5.786292428422767, eta3, (-(53.60467819078566 * eta4))))), s1, o_fma((chidiff2 * o_fma((-42.74689906831975), eta5, o_fma((-0.4600401291210382), eta3, (12.23450117663151 * eta4)))), s1, o_fma(chidiff2, o_fma((-85.19723511005235), eta5, o_fma((-3.4614418482110163), eta3, (35.464117772624164 * eta4))), ((chidiff1 * delta) * o_fma(170.9266105597438, eta5, o_fma(2.0328561081997463, eta3, (-(46.18751757691501 * eta4))))))))));
}

// hm_amp_fits_ringdown.py:198
template <class R> static inline R har_amp_32_lambda(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma(s1, o_fma(0.22358643406992756, o_fma((-915.2489655397206), eta3, o_fma(365.2485484044132, eta2, o_fma((-32.06682257944444), eta1, 0.48943645614341924))), ((0.0792473022309144 * o_fma((-3206.340850767219), eta3, o_fma(1202.174780792418, eta2, o_fma((-103.65639889587327), eta1, 1.877251717679991)))) * s1)), o_fma(chidiff2, o_fma(2657.6998831179444, eta5, o_fma(115.3689949926392, eta3, (-(1159.8596972989067 * eta4)))), o_fma((chidiff1 * delta), o_fma((-6328.182823770599), eta5, o_fma((-188.37119473865533), eta3, (2151.8731700399308 * eta4))), o_fma(49.19719473681921, eta3, o_fma((-12.259851752618998), eta2, o_fma(0.36457571743142897, eta1, 0.978510781593996))))));
}

// hm_amp_fits_ringdown.py:215
template <class R> static inline R har_amp_32_sigma(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R s1 = s;
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma(s1, o_fma((-0.012086025709597246), o_fma(17277.60594350428, eta3, o_fma((-6606.1009717965735), eta2, o_fma(600.5968613752918, eta1, (-6.230497473791485)))), (-((0.06066548829900489 * o_fma(4119.373703246675, eta3, o_fma((-1567.249168668069), eta2, o_fma(142.0346574366267, eta1, (-0.9208054306316676))))) * s1))), o_fma(chidiff2, o_fma(3592.287297392387, eta5, o_fma(209.09202210427972, eta3, (-(1769.4658099037918 * eta4)))), o_fma((chidiff1 * delta), o_fma(123.86194078913776, eta5, o_fma(144.37065005786636, eta3, (-(754.4085447486738 * eta4)))), o_fma(0.13401718687342024, eta1, 1.3353917551819414))));
}

// hm_amp_fits_ringdown.py:232
template <class R> static inline R har_amp_32_rdaux1(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.s_tot_r;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R eta6 = (eta1 * eta5);
    R s1 = s;
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma(o_fma(2048.333892310575, eta5, o_fma((-1642.3057499049708), eta4, o_fma(522.5070635563092, eta3, o_fma(4.45156488896258, eta1, (-(77.39303992494544 * eta2)))))), (1.0 / o_fma(24.249594730050312, eta2, o_fma((-9.611489164758915), eta1, 1.0))), o_fma(s1, o_fma((-0.022968727462555794), o_fma((-3572465.179268999), eta6, o_fma(2693491.7075009225, eta5, o_fma((-710313.3016672594), eta4, o_fma(77424.12614733395, eta3, o_fma(83.66854837403105, eta1, (-(3330.6261333413177 * eta2))))))), ((0.0014795114305436387 * o_fma(70606676.6311127, eta6, o_fma((-51161109.98398143), eta5, o_fma(13705532.554135624, eta4, o_fma((-1669016.9155105734), eta3, o_fma((-1672.7273629876313), eta1, (90877.38260964208 * eta2))))))) * s1)), o_fma(chidiff2, o_fma((-131.19660856923554), eta4, o_fma((-4.188795724777721), eta2, (53.39200466700963 * eta3))), ((chidiff1 * delta) * o_fma(1242.865584938088, eta4, o_fma(14.284921364132623, eta2, (-(321.26423637658746 * eta3))))))));
}

// hm_amp_fits_ringdown.py:250
template <class R> static inline R har_amp_32_rdaux2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R delta = wf.delta;
    R s = wf.chi_pn_hat;
    R chidiff = wf.dchi_half;
    R eta1 = eta;
    R eta2 = (eta1 * eta1);
    R eta3 = (eta1 * eta2);
    R eta4 = (eta1 * eta3);
    R eta5 = (eta1 * eta4);
    R eta6 = (eta1 * eta5);
    R eta7 = (eta1 * eta6);
    R s1 = s;
    R chidiff1 = chidiff;
    R chidiff2 = (chidiff1 * chidiff1);
    return o_fma(o_fma(636.0141528226893, eta5, o_fma((-693.6803787318279), eta4, o_fma(294.9496843041973, eta3, o_fma(3.902154247490771, eta1, (-(55.77521071924907 * eta2)))))), (1.0 / o_fma(19.119341007236955, eta2, o_fma((-8.56699762573719), eta1, 1.0))), o_fma(s1, o_fma((-0.16707403272774676), o_fma((-11006903.622406831), eta7, o_fma(9515538.23212259, eta6, o_fma((-3189462.4909922685), eta5, o_fma(520786.0225074669, eta4, o_fma((-41908.45179140144), eta3, o_fma(6.678916447469937, eta1, (1331.480396625797 * eta2))))))), ((0.015205286051218441 * o_fma(112013079.79786257, eta7, o_fma((-98658152.75452062), eta6, o_fma(33799252.77713386, eta5, o_fma((-5635028.227588545), eta4, o_fma(462957.5593513407, eta3, o_fma(108.10032279461095, eta1, (-(16084.215590200103 * eta2))))))))) * s1)), o_fma(chidiff2, o_fma((-440.26516625611), eta4, o_fma((-18.550171209458394), eta2, (188.99161055445936 * eta3))), ((chidiff1 * delta) * o_fma(1327.1224176812448, eta4, o_fma(13.132625215315063, eta2, (-(340.5204040505528 * eta3))))))));
}

// ---- hm_intermediate.py ----
// hm_intermediate.py:24
template <class R> static inline R hpi_phase_21_p1(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = o_fma((-7862540.0), eta6, o_fma(6825330.0, eta5, o_fma((-2288320.0), eta4, o_fma(369153.0, eta3, o_fma((-23428.1), eta2, o_fma((-1956.93), eta, (4045.84 + (7.63226 / eta))))))));
    R eq_spin = o_fma(eta2, o_fma(2689.18, s4, o_fma(2392.73, s3, o_fma((-619.837), s, (-(722.787 * s2))))), o_fma(92.6654, s4, o_fma(eta, o_fma((-1077.75), s4, o_fma(1114.51, s3, o_fma((-79.761), s, (-(355.299 * s2))))), (o_fma((-355.67), s3, o_fma((-347.273), s, (83.5428 * s2))) + (o_fma(13.6971, s3, o_fma(4.44457, s, (16.5548 * s2))) / eta)))));
    R uneq_spin = o_fma(o_fma((918.976 * wf.chi1l), o_sqrt_nonneg(o_fma((-4.0), eta, 1.0)), (-((918.976 * wf.chi2l) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))))), eta, (o_fma((91.7679 * wf.chi1l), o_sqrt_nonneg(o_fma((-4.0), eta, 1.0)), (-((91.7679 * wf.chi2l) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))))) * eta2));
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:42
template <class R> static inline R hpi_phase_21_p2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R eta4 = (eta3 * eta);
    R eta5 = (eta3 * eta2);
    R eta6 = (eta4 * eta2);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = (s3 * s);
    R no_spin = o_fma((-7862540.0), eta6, o_fma(6825330.0, eta5, o_fma((-2288320.0), eta4, o_fma(369153.0, eta3, o_fma((-27556.2), eta2, o_fma(194.72, eta, (3509.09 + (0.91868 / eta))))))));
    R eq_spin = (o_fma((eta * o_fma(2689.18, eta2, o_fma((-1077.75), eta, 92.6654))), s4, o_fma(o_fma(2392.73, eta3, o_fma((-1180.4), eta2, o_fma(117.913, eta, 5.7791))), s3, o_fma(o_fma((-619.837), eta3, o_fma(131.815, eta2, o_fma((-60.1611), eta, 0.7083999999999999))), s, (o_fma((-722.787), eta3, o_fma(278.588, eta2, o_fma((-59.2068), eta, 6.104720000000001))) * s2)))) / eta);
    R uneq_spin = ((((-91.7679) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta) * o_fma(wf.chi1l, o_fma((-1.0), eta, (-1.6012352903357276)), (wf.chi2l * o_fma(1.0, eta, 1.6012352903357276))));
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:60
template <class R> static inline R hpi_phase_21_p3(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R delta = o_sqrt_nonneg(o_fma((-4.0), eta, 1.0));
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R eta4 = (eta2 * eta2);
    R eta5 = (eta2 * eta3);
    R eta6 = (eta3 * eta3);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = o_fma((-7862540.0), eta6, o_fma(6825330.0, eta5, o_fma((-2288320.0), eta4, o_fma(369153.0, eta3, o_fma((-28651.9), eta2, o_fma(890.016, eta, 3241.68))))));
    R eq_spin = o_fma(o_fma(2689.18, eta2, o_fma((-1077.75), eta, 92.6654)), s4, o_fma(o_fma(2392.73, eta2, o_fma((-1094.59), eta, 117.913)), s3, o_fma(o_fma((-619.837), eta2, o_fma(187.641, eta, (-2.2484))), s, (o_fma((-722.787), eta2, o_fma(166.323, eta, 3.22603)) * s2))));
    R uneq_spin = (((91.7679 * wf.dchi) * delta) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:79
template <class R> static inline R hpi_phase_21_p4(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = o_fma((-7862540.0), eta6, o_fma(6825330.0, eta5, o_fma((-2288320.0), eta4, o_fma(369780.0, eta3, o_fma((-28932.5), eta2, o_fma(974.355, eta, 3160.88))))));
    R eq_spin = o_fma(o_fma((-1096.81), eta2, o_fma(577.057, eta, (-61.7446))), s4, o_fma(o_fma((-1613.0), eta2, o_fma(467.057, eta, (-20.0261))), s3, o_fma(o_fma(438.401, eta2, o_fma((-196.851), eta, 26.3355)), s, (o_fma(117.563, eta2, o_fma((-256.248), eta, 45.9957)) * s2))));
    R uneq_spin = (((65.3326 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:97
template <class R> static inline R hpi_phase_21_p5(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = o_fma(3635.76, eta3, o_fma((-1688.26), eta2, o_fma(315.911, eta, 3102.36)));
    R eq_spin = o_fma(o_fma(4689.22, eta2, o_fma((-1243.26), eta, 81.8379)), s4, o_fma(o_fma(1094.11, eta2, o_fma((-365.842), eta, 40.7054)), s3, o_fma(o_fma((-1029.76), eta2, o_fma(320.93, eta, (-23.0959))), s, (o_fma((-3079.39), eta2, o_fma(826.816, eta, (-49.5435))) * s2))));
    R uneq_spin = (((119.014 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:112
template <class R> static inline R hpi_phase_21_p6(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = o_fma((-255.245), eta3, o_fma(190.008, eta2, o_fma(4.89194, eta, 3089.18)));
    R eq_spin = o_fma(o_fma(3967.43, eta2, o_fma((-722.674), eta, (-12.8798))), s4, o_fma(o_fma(549.085, eta2, (-24.6193)), s3, o_fma(o_fma((-432.223), eta2, o_fma(57.1612, eta, 2.96997)), s, (o_fma((-2804.66), eta2, o_fma(630.516, eta, (-18.8929))) * s2))));
    R uneq_spin = (((74.0984 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:127
template <class R> static inline R hpi_phase_33_p1(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R s2 = (s * s);
    R no_spin = o_fma((-11565900.0), eta6, o_fma(10179900.0, eta5, o_fma((-3227920.0), eta4, o_fma(371303.00000000006, eta3, o_fma(18485.9, eta2, o_fma((-8727.4), eta, (4360.19 + (4.27128 / eta))))))));
    R eq_spin = (o_fma(o_fma((-34860.2), eta4, o_fma(19614.6, eta3, o_fma((-3255.6400000000003), eta2, o_fma((-251.579), eta, 11.6635)))), s, (o_fma((-74299.5), eta4, o_fma(36587.3, eta3, o_fma((-5421.92), eta2, o_fma(204.025, eta, 14.8017)))) * s2)) / eta);
    R uneq_spin = (eta * o_fma(((223.65100000000004 * wf.chi1l) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))), o_fma(1.0, eta, 3.9201300240106223), (-(((223.65100000000004 * wf.chi2l) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * o_fma(1.0, eta, 3.9201300240106223)))));
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:143
template <class R> static inline R hpi_phase_33_p2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R s2 = (s * s);
    R no_spin = o_fma((-11565900.0), eta6, o_fma(10179900.0, eta5, o_fma((-3417080.0), eta4, o_fma(518314.99999999994, eta3, o_fma((-25514.0), eta2, o_fma((-2397.09), eta, (3797.06 + (0.786684 / eta))))))));
    R eq_spin = (o_fma(o_fma((-34860.2), eta4, o_fma(19614.6, eta3, o_fma((-3520.37), eta2, o_fma(39.4668, eta, 6.7812399999999995)))), s, (o_fma((-74299.5), eta4, o_fma(36587.3, eta3, o_fma((-5914.61), eta2, o_fma(293.215, eta, 4.80384)))) * s2)) / eta);
    R uneq_spin = ((((-223.65100000000004) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta) * o_fma(wf.chi1l, o_fma((-1.0), eta, (-1.3095134830606614)), (wf.chi2l * o_fma(1.0, eta, 1.3095134830606614))));
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:159
template <class R> static inline R hpi_phase_33_p3(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R s2 = (s * s);
    R no_spin = o_fma((-11565900.0), eta6, o_fma(10179900.0, eta5, o_fma((-3525320.0), eta4, o_fma(605028.0, eta3, o_fma((-52406.1), eta2, o_fma(1796.03, eta, 3321.83))))));
    R eq_spin = o_fma(o_fma((-34860.2), eta3, o_fma(19614.6, eta2, o_fma((-3714.77), eta, 223.601))), s, (o_fma((-74299.5), eta3, o_fma(36587.3, eta2, o_fma((-5906.46), eta, 314.317))) * s2));
    R uneq_spin = (((223.651 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:175
template <class R> static inline R hpi_phase_33_p4(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R no_spin = o_fma((-1703630.0), eta6, o_fma(1172550.0, eta5, o_fma((-248477.0), eta4, o_fma(3456.2, eta3, o_fma(5139.79, eta2, o_fma((-661.15), eta, 3239.44))))));
    R eq_spin = o_fma(o_fma(67102.4, eta3, o_fma((-35337.1), eta2, o_fma(5946.15, eta, (-310.489)))), s3, o_fma(o_fma((-46537.3), eta3, o_fma(24364.0, eta2, o_fma((-4150.09), eta, 225.859))), s, (o_fma((-23603.5), eta3, o_fma(8953.98, eta2, o_fma((-994.971), eta, 35.2439))) * s2)));
    R uneq_spin = (((30.484 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:192
template <class R> static inline R hpi_phase_33_p5(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R no_spin = o_fma((-11712300.0), eta6, o_fma(9997450.0, eta5, o_fma((-3359910.0), eta4, o_fma(563997.0, eta3, o_fma((-49428.3), eta2, o_fma(2143.06, eta, 3114.3))))));
    R eq_spin = o_fma(o_fma(67102.4, eta3, o_fma((-33001.4), eta2, o_fma(5179.97, eta, (-257.524)))), s3, o_fma(o_fma((-46537.3), eta3, o_fma(23046.2, eta2, o_fma((-3705.08), eta, 190.051))), s, (o_fma((-23603.5), eta3, o_fma(10166.1, eta2, o_fma((-1414.2), eta, 63.6615))) * s2)));
    R uneq_spin = (((54.9833 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:209
template <class R> static inline R hpi_phase_33_p6(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R no_spin = o_fma((-4648580.0), eta6, o_fma(3798860.0, eta5, o_fma((-1193130.0), eta4, o_fma(179537.0, eta3, o_fma((-13003.6), eta2, o_fma(384.121, eta, 3111.46))))));
    R eq_spin = o_fma(o_fma(75501.8, eta3, o_fma((-37382.4), eta2, o_fma(5886.62, eta, (-292.855)))), s3, o_fma(o_fma((-50165.9), eta3, o_fma(24532.9, eta2, o_fma((-3834.22), eta, 182.864))), s, (o_fma((-17842.3), eta3, o_fma(6701.33, eta2, o_fma((-746.957), eta, 21.0158))) * s2)));
    R uneq_spin = (((75.5162 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:226
template <class R> static inline R hpi_phase_32_p1(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = o_fma((-3263620.0), eta6, o_fma(2863930.0, eta5, o_fma((-704442.0), eta4, o_fma((-64068.40000000001), eta3, o_fma(58234.6, eta2, o_fma((-10687.8), eta, (4414.11 + (4.21564 / eta))))))));
    R eq_spin = (o_fma((eta * o_fma(4810.61, eta2, o_fma((-1569.0299999999997), eta, 106.047))), s4, o_fma(o_fma(294.289, eta3, o_fma((-541.261), eta2, o_fma(43.8812, eta, 10.4333))), s3, o_fma(o_fma((-3970.89), eta3, o_fma(2095.72, eta2, o_fma((-610.267), eta, 6.39833))), s, (o_fma((-794.79), eta3, o_fma(331.593, eta2, o_fma((-99.1551), eta, 22.956700000000005))) * s2)))) / eta);
    R uneq_spin = (((132.244 * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta) * o_fma(wf.chi1l, o_fma((-1.0), eta, 6.227738120444028), (wf.chi2l * o_fma(1.0, eta, (-6.227738120444028)))));
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:244
template <class R> static inline R hpi_phase_32_p2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = o_fma((-3263620.0), eta6, o_fma(2863930.0, eta5, o_fma((-807629.0), eta4, o_fma(24484.2, eta3, o_fma(29218.1, eta2, o_fma((-6202.38), eta, (3980.7 + (0.956703 / eta))))))));
    R eq_spin = (o_fma((eta * o_fma(4810.61, eta2, o_fma((-1569.0299999999997), eta, 106.047))), s4, o_fma(o_fma(294.289, eta3, o_fma((-541.261), eta2, o_fma(43.8812, eta, 10.4333))), s3, o_fma(o_fma(1291.56, eta3, o_fma(75.246, eta2, o_fma((-226.825), eta, 1.92692))), s, (o_fma((-2402.94), eta3, o_fma(608.328, eta2, o_fma((-99.1551), eta, 15.328700000000001))) * s2)))) / eta);
    R uneq_spin = (((132.244 * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta) * o_fma(wf.chi1l, o_fma((-1.0), eta, 2.5769789177580837), (wf.chi2l * o_fma(1.0, eta, (-2.5769789177580837)))));
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:262
template <class R> static inline R hpi_phase_32_p3(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = o_fma((-20065000.0), eta6, o_fma(17639900.0, eta5, o_fma((-6064400.0), eta4, o_fma(1019360.0, eta3, o_fma((-84042.9), eta2, o_fma(2308.63, eta, 3416.57))))));
    R eq_spin = o_fma(o_fma(4810.61, eta2, o_fma((-1569.03), eta, 106.047)), s4, o_fma(o_fma((-114029.0), eta3, o_fma(61581.0, eta2, o_fma((-10627.4), eta, 559.705))), s3, o_fma(o_fma(12750.0, eta3, o_fma((-2582.55), eta2, o_fma((-282.354), eta, 24.6295))), s, (o_fma((-114798.0), eta3, o_fma(56407.8, eta2, o_fma((-8775.86), eta, 433.675))) * s2))));
    R uneq_spin = (((63.9466 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:280
template <class R> static inline R hpi_phase_32_p4(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = o_fma((-3263620.0), eta6, o_fma(2863930.0, eta5, o_fma((-919108.0), eta4, o_fma(127610.0, eta3, o_fma((-5980.37), eta2, o_fma((-476.909), eta, 3307.49))))));
    R eq_spin = o_fma(o_fma(4810.61, eta2, o_fma((-1569.03), eta, 106.047)), s4, o_fma(o_fma(294.289, eta2, o_fma((-370.362), eta, 43.8812)), s3, o_fma(o_fma(1291.56, eta2, o_fma((-282.354), eta, (-5.02553))), s, (o_fma((-2402.94), eta2, o_fma(740.123, eta, (-43.8823))) * s2))));
    R uneq_spin = ((((-132.244) * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:298
template <class R> static inline R hpi_phase_32_p5(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R eta7 = o_pow(eta, 7.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = o_fma((-315866000.0), eta7, o_fma(275056000.0, eta6, o_fma((-95698800.0), eta5, o_fma(17381100.0, eta4, o_fma((-1818830.0), eta3, o_fma(111203.0, eta2, o_fma((-3967.58), eta, 3259.03)))))));
    R eq_spin = o_fma(o_fma(4810.61, eta2, o_fma((-1569.03), eta, 106.047)), s4, o_fma(o_fma((-6233.99), eta2, o_fma(2183.43, eta, (-201.633))), s3, o_fma(o_fma(3810.18, eta2, o_fma((-1104.53), eta, 19.7509)), s, (o_fma((-5944.49), eta2, o_fma(2314.51, eta, (-230.07))) * s2))));
    R uneq_spin = (((112.714 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:317
template <class R> static inline R hpi_phase_32_p6(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R eta7 = o_pow(eta, 7.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = o_fma((-315866000.0), eta7, o_fma(275056000.0, eta6, o_fma((-95698800.0), eta5, o_fma(17381100.0, eta4, o_fma((-1818830.0), eta3, o_fma(111203.0, eta2, o_fma((-3967.58), eta, 3259.03)))))));
    R eq_spin = o_fma(o_fma(4810.61, eta2, o_fma((-1569.03), eta, 106.047)), s4, o_fma(o_fma((-6233.99), eta2, o_fma(2183.43, eta, (-201.633))), s3, o_fma(o_fma(3810.18, eta2, o_fma((-1104.53), eta, 19.7509)), s, (o_fma((-5944.49), eta2, o_fma(2314.51, eta, (-230.07))) * s2))));
    R uneq_spin = (((112.714 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:336
template <class R> static inline R hpi_phase_44_p1(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R no_spin = o_fma((-16080200.0), eta6, o_fma(13779200.0, eta5, o_fma((-4331970.0), eta4, o_fma(536500.0, eta3, o_fma(5534.1, eta2, o_fma((-8202.33), eta, (4349.66 + (4.34125 / eta))))))));
    R eq_spin = (o_fma((eta * o_fma((-109418.0), eta3, o_fma(57303.4, eta2, o_fma((-9508.5), eta, 492.60300000000007)))), s3, o_fma(o_fma(17900.9, eta4, o_fma((-9349.73), eta3, o_fma(1822.9100000000003, eta2, o_fma((-528.098), eta, 12.0704)))), s, (o_fma((-71523.0), eta4, o_fma(35416.6, eta3, o_fma((-5452.04), eta2, o_fma(253.334, eta, 10.4092)))) * s2))) / eta);
    R uneq_spin = ((((-262.143) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta) * o_fma(wf.chi1l, o_fma((-1.0), eta, (-3.0782778864970646)), (wf.chi2l * o_fma(1.0, eta, 3.0782778864970646))));
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:353
template <class R> static inline R hpi_phase_44_p2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R no_spin = o_fma((-16080200.0), eta6, o_fma(13779200.0, eta5, o_fma((-4501450.0), eta4, o_fma(665951.0, eta3, o_fma((-33475.8), eta2, o_fma((-2421.77), eta, (3804.19 + (0.66144 / eta))))))));
    R eq_spin = (o_fma((eta * o_fma((-109418.0), eta3, o_fma(57303.4, eta2, o_fma((-9508.5), eta, 492.60300000000007)))), s3, o_fma(o_fma(17900.9, eta4, o_fma((-7676.87), eta3, o_fma(926.576, eta2, o_fma((-172.047), eta, 5.83038)))), s, (o_fma((-71523.0), eta4, o_fma(35722.1, eta3, o_fma((-5672.02), eta2, o_fma(253.334, eta, 6.17601)))) * s2))) / eta);
    R uneq_spin = ((((-262.143) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta) * o_fma(wf.chi1l, o_fma((-1.0), eta, (-1.0543062374352932)), (wf.chi2l * o_fma(1.0, eta, 1.0543062374352932))));
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:370
template <class R> static inline R hpi_phase_44_p3(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R no_spin = o_fma((-16080200.0), eta6, o_fma(13779200.0, eta5, o_fma((-4644380.0), eta4, o_fma(777272.0, eta3, o_fma((-66340.1), eta2, o_fma(2353.58, eta, 3308.97))))));
    R eq_spin = o_fma(o_fma((-109418.0), eta3, o_fma(57303.4, eta2, o_fma((-9508.5), eta, 492.603))), s3, o_fma(o_fma(17900.9, eta3, o_fma((-7989.26), eta2, o_fma(926.576, eta, (-21.5697)))), s, (o_fma((-71523.0), eta3, o_fma(37599.5, eta2, o_fma((-6403.24), eta, 353.539))) * s2)));
    R uneq_spin = (((262.143 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:387
template <class R> static inline R hpi_phase_44_p4(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R no_spin = o_fma((-1311380.0), eta6, o_fma(908279.0, eta5, o_fma((-165169.0), eta4, o_fma((-17422.6), eta3, o_fma(8463.89, eta2, o_fma((-928.56), eta, 3245.63))))));
    R eq_spin = o_fma(o_fma(9582.58, eta3, o_fma((-4679.24), eta2, o_fma(620.695, eta, (-15.7439)))), s3, o_fma(o_fma((-6758.52), eta3, o_fma(3536.61, eta2, o_fma((-590.293), eta, 32.506))), s, (o_fma(9129.45, eta3, o_fma((-4867.87), eta2, o_fma(738.141, eta, (-25.7716)))) * s2)));
    R uneq_spin = (((87.0832 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:404
template <class R> static inline R hpi_phase_44_p5(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R eta7 = o_pow(eta, 7.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R no_spin = o_fma(228297000.00000003, eta7, o_fma((-217143000.0), eta6, o_fma(83919400.0, eta5, o_fma((-16979600.0), eta4, o_fma(1921480.0, eta3, o_fma((-119588.0), eta2, o_fma(3722.46, eta, 3108.38)))))));
    R eq_spin = o_fma(o_fma(999.417, eta2, o_fma((-356.241), eta, 20.3384)), s3, o_fma((o_fma((-529.854), eta, 118.319) * eta), s, (o_fma(516.333, eta2, o_fma((-240.648), eta, 21.0314)) * s2)));
    R uneq_spin = (((97.1364 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_intermediate.py:422
template <class R> static inline R hpi_phase_44_p6(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R no_spin = o_fma((-6011790.0), eta6, o_fma(4291930.0, eta5, o_fma((-1315230.0), eta4, o_fma(220332.0, eta3, o_fma((-20371.1), eta2, o_fma(986.752, eta, 3096.03))))));
    R eq_spin = o_fma(o_fma(9582.58, eta3, o_fma((-3070.08), eta2, o_fma(70.1931, eta, 17.0209))), s3, o_fma(o_fma((-6758.52), eta3, o_fma(2255.76, eta2, o_fma((-118.526), eta, (-9.96292)))), s, (o_fma(9129.45, eta3, o_fma((-3605.8), eta2, o_fma(370.039, eta, (-14.4869)))) * s2)));
    R uneq_spin = (((23.0759 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// ---- hm_inspiral.py ----
// hm_inspiral.py:26
template <class R> static inline R hpn_lambda_pn(R eta, int mode_tag) {
    R pi = lalpi().itself;
    R output;
    if (mode_tag == 21) {
        output = ((2.0 * pi) * o_fma((-2.0), LN2_, (-0.5)));
    } else {
        if (mode_tag == 33) {
            output = (((2.0 / 3.0) * pi) * o_fma(6.0, std::log(1.5), ((-21.0) / 5.0)));
        } else {
            if (mode_tag == 32) {
                output = (-(((2376.0 * pi) * o_fma(22.0, eta, (-5.0))) / o_fma(11880.0, eta, (-3960.0))));
            } else {
                if (mode_tag == 44) {
                    output = (((45045.0 * pi) * o_fma((320.0 * o_fma(3.0, eta, (-1.0))), LN2_, o_fma((-1193.0), eta, 336.0))) / (2.0 * o_fma(5405400.0, eta, (-1801800.0))));
                } else {
                    output = 0.0;
                }
            }
        }
    }
    return (-output);
}

// hm_inspiral.py:45
template <class R> static inline R hpn_phase_21_lambda(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R s2 = (s * s);
    R no_spin = o_fma((-133820.89317471132), eta5, o_fma(96463.68163125668, eta4, o_fma((-26847.690494515424), eta3, o_fma(3535.657736681598, eta2, o_fma((-170.08866400251395), eta, 13.664473636545068)))));
    R eq_spin = ((s * o_fma(eta2, o_fma(62149.00902591944, s2, o_fma((-7233.191207000735), s, (-35031.361998480075))), o_fma(33.8697137964237, s2, o_fma(eta, o_fma((-6961.484805326852), s2, o_fma(1198.4163078715173, s, 3642.5891077598003)), o_fma(eta3, o_fma((-149700.4915210766), s2, o_fma(16501.749243703132, s, 83493.24265292779)), o_fma((-41.55066592130464), s, 18.52571430563905)))))) / o_fma(1.0, s, 6.880288191574696));
    R uneq_spin = ((((-134.27742343186577) * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_inspiral.py:60
template <class R> static inline R hpn_phase_33_lambda(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R delta = o_sqrt_nonneg(o_fma((-4.0), eta, 1.0));
    R eta2 = (eta * eta);
    R eta3 = (eta2 * eta);
    R no_spin = o_fma(242.3428625556764, eta3, o_fma((-103.92956504520747), eta2, o_fma(9.772510519809892, eta, 4.1138398568400705)));
    R eq_spin = ((o_fma((-105.09339163109497), eta2, o_fma(26.644159828590055, eta, (-0.13253553909611435))) * s) / o_fma(0.11322426762297967, s, 1.0));
    R uneq_spin = ((((-19.705359163581168) * wf.dchi) * eta2) * delta);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_inspiral.py:73
template <class R> static inline R hpn_phase_32_lambda(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = (o_fma((-6244.001932443913), eta4, o_fma(2671.7813055097877, eta3, o_fma((-574.8672384388947), eta2, o_fma(18.424900617803107, eta, 9.913819875501506)))) / o_fma((-0.9103118343073325), eta, 1.0));
    R eq_spin = o_fma(o_fma(42379.205045791656, eta3, o_fma((-18636.271414571278), eta2, o_fma(2009.7361967331492, eta, 1.8403161863444328))), s4, o_fma(o_fma((-273.23750933544176), eta3, o_fma(708.1455154504333, eta2, o_fma((-138.36970336658558), eta, (-1.28324755577382)))), s3, o_fma(o_fma(5894.355429022858, eta3, o_fma((-2233.9319708029775), eta2, o_fma(245.06757304950986, eta, (-4.367632806613781)))), s, (o_fma((-40928.07304790013), eta3, o_fma(17608.172965575013, eta2, o_fma((-1876.760129419146), eta, (-1.375112297530783)))) * s2))));
    R uneq_spin = (((wf.dchi * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2) * o_fma((2155.472229664981 * eta), s, o_fma((1566.1242344157668 * wf.chi2l), eta, o_fma((-(1566.1242344157668 * wf.chi1l)), eta, (-105.34550407768225)))));
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_inspiral.py:89
template <class R> static inline R hpn_phase_44_lambda(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R s2 = (s * s);
    R no_spin = o_fma((-1538.6661348106031), eta5, o_fma(1685.5912722190276, eta4, o_fma((-1162.954360723399), eta3, o_fma(160.43721442910618, eta2, o_fma((-21.277760168559862), eta, 5.254484747463392)))));
    R eq_spin = o_fma(o_fma((-810.7773268493444), eta3, o_fma(246.8787141453734, eta2, o_fma((-10.945895160727437), eta, 0.007067861615983771))), s, (o_fma(621.6920322846844, eta3, o_fma((-176.4987316167203), eta2, o_fma(4.530539154777984, eta, 0.17447830920234977))) * s2));
    R uneq_spin = ((((-8.384066369867833) * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// ---- hm_ringdown.py ----
// hm_ringdown.py:23
template <class R> static inline R hpr_phase_22_alpha2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R delta = o_sqrt_nonneg(o_fma((-4.0), eta, 1.0));
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R no_spin = o_fma(55.45508989446867, eta4, o_fma((-31.330215053905395), eta3, o_fma(6.510807976353186, eta2, o_fma((-0.37138987533788487), eta, 0.2088669311744758))));
    R eq_spin = ((o_fma(38.61300158832203, eta3, o_fma((-16.874355161681766), eta2, o_fma(1.6966740823756759, eta, 0.2393965714370633))) * s) / o_fma((-0.633218538432246), s, 1.0));
    R uneq_spin = ((wf.dchi * o_fma(0.9088578269496244, o_pow(eta, 2.5), ((15.619592332008951 * wf.dchi) * o_pow(eta, 3.5)))) * delta);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_ringdown.py:37
template <class R> static inline R hpr_phase_22_alpha_l(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R delta = o_sqrt_nonneg(o_fma((-4.0), eta, 1.0));
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R s2 = (s * s);
    R no_spin = (eta * o_fma(27.623649807617376, eta3, o_fma((-16.504334734464244), eta2, o_fma(2.5400257699690143, eta, (-1.1926122248825484)))));
    R eq_spin = o_fma((eta4 * s), o_fma(180.8345521274853, s2, o_fma((-39.83460092417137), s, (-72.50310678862684))), o_fma((eta2 * s), o_fma(8.170616575493503, s2, o_fma(0.8743576195364632, s, (-4.7282481007397825))), o_fma((eta3 * s), o_fma((-77.2346297158916), s2, o_fma(9.700178927988006, s, 35.803988443700824)), ((eta * s) * o_fma((-0.06417449517826644), s2, o_fma((-0.21477847929548569), s, 0.1034526554654983))))));
    R uneq_spin = (o_fma((17.588573345324154 * (wf.chi2l * wf.chi2l)), o_pow(eta, 4.5), o_fma((-((35.17714669064831 * wf.chi1l) * wf.chi2l)), o_pow(eta, 4.5), o_fma((17.588573345324154 * (wf.chi1l * wf.chi1l)), o_pow(eta, 4.5), o_fma(((-0.7428134042821221) * wf.chi1l), o_pow(eta, 3.5), ((0.7428134042821221 * wf.chi2l) * o_pow(eta, 3.5)))))) * delta);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_ringdown.py:52
template <class R> static inline R hpr_phase_32_spheroidal_time_shift(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = o_fma(739270.8814129328, eta5, o_fma((-364129.24735853914), eta4, o_fma(61559.132976189896, eta3, o_fma((-4565.033758777737), eta2, o_fma(167.95086712701223, eta, 11.851438981981772)))));
    R eq_spin = o_fma(o_fma((-65993.84497923202), eta3, o_fma(24793.809334683185, eta2, o_fma((-2292.778112523392), eta, 15.585154698977842))), s4, o_fma(o_fma((-60929.87465551446), eta3, o_fma(24145.932943269272, eta2, o_fma((-2244.4613237912527), eta, 3.4881615575084797))), s3, o_fma(o_fma(26929.677144312944, eta3, o_fma((-8046.364492927503), eta2, o_fma(434.31707030999445, eta, 9.506768471271634))), s, (o_fma(3575.347142399199, eta3, o_fma(1334.1062451631644, eta2, o_fma((-307.67253970367034), eta, (-5.949655484033632)))) * s2))));
    R uneq_spin = (((465.7904934097202 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_ringdown.py:69
template <class R> static inline R hpr_phase_32_spheroidal_phase_shift(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R eta6 = o_pow(eta, 6.0);
    R eta7 = o_pow(eta, 7.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = o_fma((-7612233.821752137), eta7, o_fma(5614617.173188322, eta6, o_fma((-1631296.8467540336), eta5, o_fma(246313.12887984765, eta4, o_fma((-21256.376324666326), eta3, o_fma(1056.2426481245027, eta2, o_fma((-22.209549522908667), eta, (-1.3328895897490733))))))));
    R eq_spin = ((s * o_fma(eta3, o_fma(805.0266029288334, s4, o_fma((-1493.177074873678), s3, o_fma((-2683.947280170815), s2, o_fma((-22.59781899512552), s, 62.632493533037625)))), o_fma(eta, o_fma(24.02858084544326, s4, o_fma((-34.498125910065156), s3, o_fma((-97.16014978329918), s2, o_fma((-35.074072557604445), s, (-47.475568056234245))))), o_fma(0.7638287749489343, s4, o_fma(eta2, o_fma((-312.1010566468068), s4, o_fma(546.6246807461155, s3, o_fma(1086.7504825459278, s2, o_fma(135.25216875394943, s, 196.7055281997748)))), o_fma((-1.9586085340429995), s3, o_fma((-1.1239505323267036), s2, o_fma(0.9960210841611344, s, (-1.622727240110213))))))))) / o_fma(1.0, s, (-2.950271397057221)));
    R uneq_spin = ((o_sqrt_nonneg(o_fma((-4.0), eta, 1.0)) * o_fma((wf.chi1l * o_pow(eta, 2.5)), o_fma(30.01812659282717, s, (-88.56162028006072)), o_fma((wf.chi1l * eta2), o_fma(14.617728550838805, s, (-43.126266433486435)), o_fma((wf.chi2l * o_pow(eta, 2.5)), o_fma((-30.01812659282717), s, 88.56162028006072), ((wf.chi2l * eta2) * o_fma((-14.617728550838805), s, 43.126266433486435)))))) / o_fma(1.0, s, (-2.950271397057221)));
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_ringdown.py:88
template <class R> static inline R hpr_phase_32_p1(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R eta4 = o_pow(eta, 4.0);
    R eta5 = o_pow(eta, 5.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R s5 = o_pow(s, 5.0);
    R no_spin = o_fma(1567405.3633767858, eta5, o_fma((-817182.2896823225), eta4, o_fma(149846.7281073725, eta3, o_fma((-12569.748101922158), eta2, o_fma(426.8372805022653, eta, 3169.372056189274)))));
    R eq_spin = o_fma(o_fma((-384336.57477596396), eta3, o_fma(162941.8009556697, eta2, o_fma((-16556.65370439302), eta, (-61.01931672442576)))), s5, o_fma(o_fma((-139322.02687945773), eta3, o_fma(53893.9479532688, eta2, o_fma((-4901.068176970293), eta, (-24.72321125342808)))), s4, o_fma(o_fma(201204.5096556517, eta3, o_fma((-91664.12621864353), eta2, o_fma(10145.53174542332, eta, 13.918237996338371))), s3, o_fma(o_fma((-3785.49764771212), eta3, o_fma(7855.316419853637, eta2, o_fma((-1762.6573670619173), eta, 19.23408352151287))), s, (o_fma(20497.5021807654, eta3, o_fma((-5615.908682338113), eta2, o_fma(336.8340966473415, eta, (-42.88446003698396)))) * s2)))));
    R uneq_spin = (((wf.dchi * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2) * o_fma((13275.623692212472 * eta), s, o_fma((1600.240100295189 * wf.chi2l), eta, o_fma((-(1600.240100295189 * wf.chi1l)), eta, 641.2473192044652))));
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_ringdown.py:106
template <class R> static inline R hpr_phase_32_p2(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = o_fma(7475.062269742079, eta3, o_fma((-2636.4344627081873), eta2, o_fma(206.09687819102305, eta, 3131.0260952676376)));
    R eq_spin = o_fma(o_fma(69666.01833764123, eta3, o_fma((-34444.2579678986), eta2, o_fma(5119.138772157134, eta, (-227.8023564332453)))), s4, o_fma(o_fma((-27654.067482558712), eta3, o_fma(7269.262546204149, eta2, o_fma(429.0787542102207, eta, (-132.49105074500454)))), s3, o_fma(o_fma(10514.68111669422, eta3, o_fma((-434.60154548208334), eta2, o_fma((-691.9815135740145), eta, 49.90874152040307))), s, (o_fma((-56142.13736008524), eta3, o_fma(26748.805404989867, eta2, o_fma((-3458.2579971189534), eta, 97.3078084654917))) * s2))));
    R uneq_spin = (((477.51566939885424 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_ringdown.py:121
template <class R> static inline R hpr_phase_32_p3(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R eta3 = o_pow(eta, 3.0);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = o_fma(977.6115755788503, eta3, o_fma((-586.2469821978381), eta2, o_fma(76.94679795837645, eta, 3082.803556599222)));
    R eq_spin = o_fma(o_fma(69666.01833764123, eta3, o_fma((-34444.2579678986), eta2, o_fma(5119.138772157134, eta, (-227.8023564332453)))), s4, o_fma(o_fma(16259.911572457446, eta3, o_fma((-10971.864789147161), eta2, o_fma(2268.4693587493093, eta, (-144.96370424517866)))), s3, o_fma(o_fma(2472.6476419567534, eta3, o_fma(1775.4343704616288, eta2, o_fma((-807.7353772747749), eta, 45.08944710349874))), s, (o_fma((-25583.314298758105), eta3, o_fma(13821.251641893134, eta2, o_fma((-2224.9613131172046), eta, 95.57355060136699))) * s2))));
    R uneq_spin = (((378.2359918274837 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// hm_ringdown.py:136
template <class R> static inline R hpr_phase_32_p4(const WfVars<R>& wf) {
    R eta = wf.eta;
    R s = wf.s_tot_r;
    R eta2 = (eta * eta);
    R s2 = (s * s);
    R s3 = o_pow(s, 3.0);
    R s4 = o_pow(s, 4.0);
    R no_spin = o_fma((-357.38692756785395), eta2, o_fma(64.99844502520415, eta, 3077.0657367004565));
    R eq_spin = o_fma(o_fma((-41414.05689348732), eta2, o_fma(95314.99988114933, o_pow(eta, 3.0), o_fma(4900.706039920717, eta, (-169.30909412804587)))), s4, o_fma(o_fma((-26766.730897942085), eta2, o_fma(62863.79877094988, o_pow(eta, 3.0), o_fma(3169.3397351803583, eta, (-126.02362949830213)))), s3, o_fma(o_fma(5700.682624203565, eta2, o_fma((-9490.641676924794), o_pow(eta, 3.0), o_fma((-986.7751755509875), eta, 34.793450080444714))), s, (o_fma(11008.881935880598, eta2, o_fma((-19906.416384606226), o_pow(eta, 3.0), o_fma((-1644.6690499868596), eta, 57.38106384558743))) * s2))));
    R uneq_spin = (((390.5443469721231 * wf.dchi) * o_sqrt_nonneg(o_fma((-4.0), eta, 1.0))) * eta2);
    return ((no_spin + eq_spin) + uneq_spin);
}

// xhm/reference.py:189 (numpy.linalg.inv in float64)
static const double hm_inspiral_inverse[3][3] = {{193.23732329876478, -147.86916939486258, 38.22764042611204}, {-405.9225177333149, 334.17264854592065, -90.23877138864589}, {212.68519443455014, -186.30347915105807, 53.01113096253385}};
