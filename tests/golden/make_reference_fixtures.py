"""Writes tests/golden/reference_tests.json: the input systems and golden values
of the reference's own test suite, transcribed as data.

Source of every number: platforms/reference/tests/TestReferenceMPIDForce.cpp
  * make_waterbox (lines 109-660), 6-atom dimer coordinates (lines 226-236)
  * make_methanolbox (lines 662-880), 12-atom coordinates (lines 845-860)
  * make_charge_square (lines 64-99)
  * expected energies/forces: testMethanolDimer* (972-1299), testWaterDimer* (1301-1601),
    test14Scaling* (1603-1700).
The same values appear in platforms/cuda/tests/TestCudaMPIDForce.cpp.
Coordinates in the reference are Angstrom and multiplied by NmPerAngstrom (0.1).

Run:  python tests/golden/make_reference_fixtures.py
"""
import json
import os

NM = 0.1


def water_dimer():
    # per-type parameters (TestReferenceMPIDForce.cpp:125-190)
    types = {
        "O": dict(charge=-0.51966, dipole=[0.0, 0.0, 0.00755612136146],
                  quadrupole=[0.000354030721139, 0.0, -0.000390257077096, 0.0, 0.0, 3.62263559571e-05],
                  octopole=[0, 0, 0, 0, -6.285758282686837e-07, 0, -9.452653225954594e-08, 0, 0, 7.231018665791977e-07],
                  axis=1, anchors=[1, 2, 0], alphas=[0.000837] * 3, thole=0.39,
                  cov12=[1, 2], cov13=[]),
        "H1": dict(charge=0.25983, dipole=[-0.00204209484795, 0.0, -0.00307875299958],
                   quadrupole=[-3.42848248983e-05, 0.0, -0.000100240875193, -1.89485963908e-06, 0.0, 0.000134525700091],
                   octopole=[-2.405600937552608e-07, 0, -6.415084018183151e-08, 0, -1.152422607026746e-06, 0,
                             -2.558537436767218e-06, 3.047102424084479e-07, 0, 3.710960043793964e-06],
                   axis=0, anchors=[-1, 1, 0], alphas=[0.000496] * 3, thole=0.39,
                   cov12=[-1], cov13=[1]),
        "H2": dict(charge=0.25983, dipole=[-0.00204209484795, 0.0, -0.00307875299958],
                   quadrupole=[-3.42848248983e-05, 0.0, -0.000100240875193, -1.89485963908e-06, 0.0, 0.000134525700091],
                   octopole=[-2.405600937552608e-07, 0, -6.415084018183151e-08, 0, -1.152422607026746e-06, 0,
                             -2.558537436767218e-06, 3.047102424084479e-07, 0, 3.710960043793964e-06],
                   axis=0, anchors=[-2, -1, 0], alphas=[0.000496] * 3, thole=0.39,
                   cov12=[-2], cov13=[-1]),
    }
    coords = [[2.0, 2.0, 2.0], [2.5, 2.0, 3.0], [1.5, 2.0, 3.0],
              [0.0, 0.0, 0.0], [0.5, 0.0, 1.0], [-0.5, 0.0, 1.0]]
    order = ["O", "H1", "H2"]
    atoms = []
    for a in range(6):
        t = types[order[a % 3]]
        z = a + t["anchors"][0]
        x = a + t["anchors"][1]
        y = -1 if t["anchors"][2] == 0 else a + t["anchors"][2]
        atoms.append(dict(charge=t["charge"], dipole=t["dipole"], quadrupole=t["quadrupole"],
                          octopole=t["octopole"], axisType=t["axis"], atomZ=z, atomX=x, atomY=y,
                          thole=t["thole"], alphas=t["alphas"],
                          cov12=[a + c for c in t["cov12"]], cov13=[a + c for c in t["cov13"]], cov14=[]))
    return dict(atoms=atoms, positions=[[c * NM for c in p] for p in coords])


def methanol_dimer():
    zero3, zero6, zero10 = [0.0] * 3, [0.0] * 6, [0.0] * 10
    types = {
        "C1": dict(charge=-0.140, dipole=zero3, quadrupole=zero6, octopole=zero10, axis=4,
                   anchors=[1, 0, 0], alphas=[0.001] * 3, thole=1.3, cov12=[1, 3, 4, 5], cov13=[2]),
        "O1": dict(charge=-0.460, dipole=[0.00026405942708641, 0.0, 0.00550661803258754],
                   quadrupole=[9.383755641232907e-05, 0, -0.0001547997648007625, -1.577493985246555e-06, 0,
                               6.096220838843343e-05],
                   octopole=[-3.230426667733178e-08, 0, 3.684859776955582e-08, 0, -2.245492298396793e-07, 0,
                             7.675967953604524e-07, -4.445541285871346e-09, 0, -5.43047565520773e-07],
                   axis=0, anchors=[-1, 1, 0], alphas=[0.00100024, 0.00125025, 0.00083350], thole=1.3,
                   cov12=[-1, 1], cov13=[2, 3, 4]),
        "HO1": dict(charge=0.360, dipole=zero3, quadrupole=zero6, octopole=zero10, axis=4,
                    anchors=[-1, 0, 0], alphas=zero3, thole=0.0, cov12=[-1], cov13=[-1]),
        "H1A": dict(charge=0.080, dipole=zero3, quadrupole=zero6, octopole=zero10, axis=4,
                    anchors=[-3, 0, 0], alphas=zero3, thole=0.0, cov12=[-3], cov13=[-2, 1, 2]),
        "H1B": dict(charge=0.080, dipole=zero3, quadrupole=zero6, octopole=zero10, axis=4,
                    anchors=[-4, 0, 0], alphas=zero3, thole=0.0, cov12=[-4], cov13=[-3, -1, 1]),
        "H1C": dict(charge=0.080, dipole=zero3, quadrupole=zero6, octopole=zero10, axis=4,
                    anchors=[-5, 0, 0], alphas=zero3, thole=0.0, cov12=[-5], cov13=[-4, -2, -1]),
    }
    coords = [
        [1.6118739816, -7.7986654421, -9.3388011053], [0.4344388195, -8.6290855266, -9.4591523136],
        [-0.2932869802, -8.1452383606, -9.0381926002], [1.5101797393, -6.7361319725, -9.6470249020],
        [1.8800020341, -7.7312323778, -8.2627522535], [2.3828031354, -8.2685172700, -9.9862796759],
        [-2.3016642008, -3.3801483374, -4.5239842701], [-2.6774345292, -3.8370280231, -3.2318499504],
        [-1.9568218092, -3.4707595618, -2.6956925837], [-1.4748236015, -3.9573461155, -4.9903514535],
        [-3.2561339708, -3.3690912389, -5.0924789477], [-1.9925289806, -2.3413378186, -4.2797935219]]
    order = ["C1", "O1", "HO1", "H1A", "H1B", "H1C"]
    atoms = []
    for a in range(12):
        t = types[order[a % 6]]
        z = a + t["anchors"][0]
        x = -1 if t["anchors"][1] == 0 else a + t["anchors"][1]
        y = -1 if t["anchors"][2] == 0 else a + t["anchors"][2]
        atoms.append(dict(charge=t["charge"], dipole=t["dipole"], quadrupole=t["quadrupole"],
                          octopole=t["octopole"], axisType=t["axis"], atomZ=z, atomX=x, atomY=y,
                          thole=t["thole"], alphas=t["alphas"],
                          cov12=[a + c for c in t["cov12"]], cov13=[a + c for c in t["cov13"]], cov14=[]))
    return dict(atoms=atoms, positions=[[c * NM for c in p] for p in coords])


def charge_square():
    pos = [[0.1, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.1, 0.0], [0.1, 0.1, 0.0]]
    q = [1.0, 1.0, -1.0, -1.0]
    cov = [([1], [2], [3]), ([0, 2], [3], []), ([1, 3], [0], []), ([2], [1], [0])]
    atoms = []
    for a in range(4):
        atoms.append(dict(charge=q[a], dipole=[0.0] * 3, quadrupole=[0.0] * 6, octopole=[0.0] * 10,
                          axisType=5, atomZ=0, atomX=0, atomY=0, thole=0.0, alphas=[0.0] * 3,
                          cov12=cov[a][0], cov13=cov[a][1], cov14=cov[a][2]))
    return dict(atoms=atoms, positions=pos)


def main():
    water = water_dimer()
    methanol = methanol_dimer()
    square = charge_square()
    cases = []
    # methanol PME (TestReferenceMPIDForce.cpp:972-1027, 1081-1137)
    mpme = dict(system="methanol", method=1, alpha=4.5, grid=64, cutoff=1.2, box=2.461817,
                default_thole=3.0)
    cases.append(dict(name="methanol_pme_direct", polarization=1, target_eps=1e-5, **mpme,
                      energy=100.048119, source="TestReferenceMPIDForce.cpp:1003-1016",
                      forces=[[0.4407512632, 0.9533272891, 0.2662227116], [2.777987186, 1.840858052, -1.333640734],
                              [-171.6344629, -55.43492185, 12.32086698], [54.15942246, 41.78995026, -18.66314014],
                              [67.2048408, 12.09783528, 23.99010209], [46.52456681, -2.469100554, -16.59349719],
                              [0.9792294678, -1.511490917, -2.188302524], [2.487119928, -3.695492757, -6.600407624],
                              [12.53120547, -46.68073149, 204.8161], [13.32630506, -12.47148185, -64.76480367],
                              [-26.51166625, 3.163119144, -46.15867739], [-2.28426176, 62.41964403, -85.09330594]]))
    cases.append(dict(name="methanol_pme_mutual", polarization=0, target_eps=1e-9, **mpme,
                      energy=100.0480699, source="TestReferenceMPIDForce.cpp:1113-1126",
                      forces=[[0.4388490535, 0.9521698628, 0.2671848666], [2.77743403, 1.838052623, -1.324369675],
                              [-171.6330632, -55.43091584, 12.31209076], [54.15929787, 41.79141719, -18.66382841],
                              [67.20544012, 12.09872923, 23.98850965], [46.52468443, -2.468565193, -16.59413613],
                              [0.9812786184, -1.522494372, -2.19205581], [2.482142255, -3.688423769, -6.596946686],
                              [12.53264302, -46.6801623, 204.8169521], [13.32737364, -12.47126727, -64.76450554],
                              [-26.51129625, 3.16340949, -46.15822949], [-2.283744668, 62.41956649, -85.09314563]]))
    # Extrapolated (OPT) polarization with the default coefficients
    # (-0.154, 0.017, 0.658, 0.474; MPIDForce.cpp:46-49),
    # TestReferenceMPIDForce.cpp:1192-1245
    cases.append(dict(name="methanol_pme_extrapolated", polarization=2, target_eps=1e-5, **mpme,
                      energy=100.0480906, source="TestReferenceMPIDForce.cpp:1223-1235",
                      forces=[[0.438635498, 0.9521799034, 0.2673223945], [2.777143984, 1.837985034, -1.32278651],
                              [-171.6327822, -55.43082349, 12.31056078], [54.15935455, 41.79153141, -18.66398402],
                              [67.20558687, 12.0987565, 23.9882874], [46.52473894, -2.46854161, -16.59424569],
                              [0.9814929295, -1.523627818, -2.192349756], [2.481269398, -3.686996821, -6.596448702],
                              [12.53309724, -46.68066444, 204.8169435], [13.32746646, -12.47126075, -64.76446756],
                              [-26.51126553, 3.163426109, -46.15818436], [-2.283699129, 62.41955233, -85.09312721]]))
    # water PME (TestReferenceMPIDForce.cpp:1350-1400, 1449-1501)
    wpme = dict(system="water", method=1, alpha=3.0, grid=64, cutoff=0.6, box=2.0, default_thole=3.0)
    cases.append(dict(name="water_pme_direct", polarization=1, target_eps=1e-5, **wpme,
                      energy=-2.523318862, source="TestReferenceMPIDForce.cpp:1381-1388",
                      forces=[[-138.9578383, -183.3187212, 31.05996292], [36.78883138, -5.591080652, 7.601999899],
                              [41.46403045, 118.9693325, 34.16137849], [-116.5222458, -100.9480058, -32.82501978],
                              [126.6226866, 166.1966239, -17.03839082], [50.6045891, 4.691845173, -22.96008637]]))
    cases.append(dict(name="water_pme_mutual", polarization=0, target_eps=1e-8, **wpme,
                      energy=-2.533082539, source="TestReferenceMPIDForce.cpp:1482-1489",
                      forces=[[-140.0801113, -184.8502938, 30.90206227], [37.10990648, -5.575145037, 7.692659824],
                              [41.55181662, 119.50422, 34.31895915], [-117.0338412, -101.5516429, -33.10846732],
                              [127.7947382, 167.7834207, -16.87461724], [50.65754476, 4.689434353, -22.93075376]]))
    # TestReferenceMPIDForce.cpp:1503-1550
    cases.append(dict(name="water_pme_extrapolated", polarization=2, target_eps=1e-5, **wpme,
                      energy=-2.527846018, source="TestReferenceMPIDForce.cpp:1535-1541",
                      forces=[[-140.156749, -184.9802098, 30.95576142], [37.14821556, -5.561204369, 7.691990976],
                              [41.56730743, 119.5695275, 34.32030003], [-117.1007878, -101.6308286, -33.11577527],
                              [127.9092349, 167.929753, -16.90691988], [50.63283256, 4.672955505, -22.94551458]]))
    # NoCutoff (TestReferenceMPIDForce.cpp:1029-1080, 1139-1190, 1249-1299,
    # 1301-1347, 1401-1447, 1553-1601): every pair, no periodic images; the box
    # is given to the context but unused; default Thole width 3.0
    mnc = dict(system="methanol", method=0, alpha=0.0, grid=0, cutoff=1.0, box=2.461817, default_thole=3.0)
    wnc = dict(system="water", method=0, alpha=0.0, grid=0, cutoff=1.0, box=2.0, default_thole=3.0)
    cases.append(dict(name="methanol_nocut_direct", polarization=1, target_eps=1e-5, **mnc,
                      energy=100.1426571, source="TestReferenceMPIDForce.cpp:1055-1068",
                      forces=[[0.6174848862, 1.244067416, 0.348981665], [3.137044182, 3.021402996, -1.106973251],
                              [-171.9428202, -56.29908471, 12.15602242], [54.08049294, 41.61038079, -18.72605327],
                              [67.10967846, 11.90263439, 23.93203305], [46.43698088, -2.657546533, -16.65203133],
                              [1.1456374, -1.273419433, -2.195068492], [3.279289766, -2.747123834, -6.42588385],
                              [11.97737104, -47.37054551, 204.7348033], [13.21305881, -12.64877074, -64.79230192],
                              [-26.64547941, 2.986613971, -46.17381233], [-2.408738783, 62.23139119, -85.099716]]))
    cases.append(dict(name="methanol_nocut_mutual", polarization=0, target_eps=1e-9, **mnc,
                      energy=100.1424251, source="TestReferenceMPIDForce.cpp:1166-1179",
                      forces=[[0.6146835636, 1.24282615, 0.3501711642], [3.13797926, 3.015595591, -1.097188673],
                              [-171.9420969, -56.29281618, 12.14742654], [54.08018498, 41.61211644, -18.72680649],
                              [67.11010165, 11.90386444, 23.93032756], [46.43703861, -2.656708806, -16.65281219],
                              [1.148286492, -1.284971833, -2.197342953], [3.272556789, -2.742411944, -6.424646877],
                              [11.97974305, -47.36853963, 204.7356438], [13.21447088, -12.64805524, -64.79191248],
                              [-26.64493967, 2.987395526, -46.17328869], [-2.40800871, 62.23170548, -85.09957069]]))
    cases.append(dict(name="methanol_nocut_extrapolated", polarization=2, target_eps=1e-5, **mnc,
                      energy=100.1424271, source="TestReferenceMPIDForce.cpp:1275-1288",
                      forces=[[0.6144183532, 1.242877198, 0.3502970714], [3.13789228, 3.015096318, -1.095457146],
                              [-171.941967, -56.29241757, 12.14584015], [54.08021516, 41.61225418, -18.72696219],
                              [67.11022091, 11.90392367, 23.93009922], [46.43708221, -2.656656187, -16.65293107],
                              [1.148617423, -1.286157531, -2.197592114], [3.271428661, -2.741098117, -6.424340583],
                              [11.98035816, -47.36903306, 204.7356612], [13.21458611, -12.64799137, -64.79184765],
                              [-26.64490265, 2.987465959, -46.17322487], [-2.40794957, 62.2317365, -85.09954205]]))
    cases.append(dict(name="water_nocut_direct", polarization=1, target_eps=1e-5, **wnc,
                      energy=-1.949902453, source="TestReferenceMPIDForce.cpp:1329-1336",
                      forces=[[-138.7310812, -182.9709838, 35.70961618], [37.1153441, -5.548490702, 5.04277195],
                              [41.13860764, 118.8270727, 31.47279046], [-116.4297925, -100.864177, -27.61517965],
                              [126.6370205, 165.8966158, -19.33373258], [50.26990146, 4.659962993, -25.27626636]]))
    cases.append(dict(name="water_nocut_mutual", polarization=0, target_eps=1e-8, **wnc,
                      energy=-1.952917117, source="TestReferenceMPIDForce.cpp:1430-1437",
                      forces=[[-139.7835608, -184.4337529, 35.62953533], [37.434981, -5.522902943, 5.11681405],
                              [41.23101208, 119.3674074, 31.61700973], [-116.9476192, -101.4714619, -27.86430037],
                              [127.7709383, 167.4188741, -19.23010497], [50.29424862, 4.641836191, -25.26895377]]))
    cases.append(dict(name="water_nocut_extrapolated", polarization=2, target_eps=1e-5, **wnc,
                      energy=-1.94668563, source="TestReferenceMPIDForce.cpp:1583-1590",
                      forces=[[-139.8529084, -184.5568497, 35.69566243], [37.47391144, -5.507167568, 5.113662284],
                              [41.24807762, 119.4349691, 31.61668322], [-117.0172298, -101.5532625, -27.86681578],
                              [127.8817331, 167.5584812, -19.27180201], [50.26641605, 4.623829403, -25.28739015]]))
    # 1-4 scaling without cutoff (TestReferenceMPIDForce.cpp:1603-1646): closed form, tolerance 1e-5
    for s14, e in ((1.0, -1389.35), (0.5, -1389.35 / 2), (0.0, 0.0)):
        cases.append(dict(name=f"charge_square_nocut_s14_{s14}", system="square", method=0, alpha=0.0, grid=0,
                          cutoff=1.0, box=2.0, default_thole=5.0, polarization=1, target_eps=1e-5,
                          scale14=s14, energy=e, energy_tol=1e-5, source="TestReferenceMPIDForce.cpp:1603-1646"))
    # 1-4 scaling with PME (TestReferenceMPIDForce.cpp:1648-1700): closed form, tolerance 1e-3
    for s14, e in ((1.0, -1389.35), (0.5, -1389.35 / 2), (0.0, 0.0)):
        cases.append(dict(name=f"charge_square_pme_s14_{s14}", system="square", method=1, alpha=0.001, grid=64,
                          cutoff=0.4, box=4.0, default_thole=5.0, polarization=1, target_eps=1e-5,
                          scale14=s14, energy=e, energy_tol=1e-3, source="TestReferenceMPIDForce.cpp:1648-1700"))
    out = dict(systems=dict(water=water, methanol=methanol, square=square), cases=cases, tolerance=1e-4)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_tests.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
