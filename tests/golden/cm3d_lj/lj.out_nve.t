Reading in set from file "lj.set"
Reading in parameter from file lj.top
There are 0 groups and 1 types of atoms
Searching intermolecular data base data_base for interactions
There were 1 atom types and 1 interations (out of 1)
Long range correction factor V = -35090.1 K A^3, W = 70084.3 K A^3
There were 0 bonds types and 0 bonds
There were 0 types and 0 bends
Number of torss types = 0, Number of total torss = 0
Number of harmonic = 0, Number of power series = 0
# of one-fours = 0, and n14_types = 0
scale_onefour = 1 and scale_onefour_e = 1
number of bondx types = 0 number of total cross bonds = 0
There are 0 charged particles (out of a total of 256)
There were 0 k-vectors set up.
Self correction term = 0 K
Back ground term = 0 K = 0 e
Allocating exclusion list memory
Excluding bonded interactions
There were 0 interaction exclusions found
Setting up electrostatic exclusion list
Setting up electrostatic exclusion list (ncharge = 0)
There are 0 real space electrostatic corrections
There are 0 coulomb space electrostatic corrections
Barostat set up: mvol = 255 and you have 0 barostates
Doing cold start from initial file "lj.init"
256 Atoms read in with a Cell Volume of 301.177
***********************************
* istart=1 or 2 : Reset averages  *
*      and counters and deriv.    *
***********************************
Calculating initial forces
Setting up link list neighbors (~343 cells/processor)
Number of interacting cells (procs #0) is 47334 (out of 58996)
Number of electrostatic cells (procs #0) is 58996 (out of 58996)
Counting vdw pairs
Van der Waals Neighbor list memory changed!
Electrostatic Neighbor list memory changed!
Calculating initial potential energies

type of calculation   = 1
units                 = 0
periodicity           = 3
# of atoms            = 256
# of frozen atoms     = 0
MD time step          = 0.001 ps
# of inter time steps = 1 (dt = 0.001)
# of Torsion steps    = 1 (dt = 0.001)
# of Bond/Bend steps  = 1 (dt = 0.001)
# of thermostats      = 0
# of nh chains links  = 0
# of barostats        = 0
Temperature is not thermostated
Volume will not fluctuate.

Averages reported every       100 ts
Restart info dumped every     500 ts
A config is save every        10 ts
A velocity file is save every 10 ts
A force file is save every    0 ts
Instant quantaties recorded   100 ts

Calcutaion time step 1 to 1000.
total time of simulation 1 ps.


  *** Initial Report ***
npairs = s=2334 t=7851 se=0 te=0
  density  = 1.17356 g/cc
  Volume   = 301.177 A^3 (cbrt = 6.70307)
  Temp     = 1 K
  Pressure = 3.97673 K/A^3 = 541.857 ATM

nprs=2334, nprt=7851, nprse=0, nprte=0, <ts/update>=0
<Uer/N>  	=        0.000000	~       -5.268099 K
<Ke /N>  	=        0.000000	~        1.494141 K
<inter>  	=        0.000000	~    -1232.123279 K
<Ham/N>  	=        0.000000	~       -3.323263 K
<temp>   	=        0.000000	~        1.000000 K
<pres>   	=        0.000000	~        3.976735 K/A^3
<pres>   	=        0.000000	~      541.857159 atm
Pressure Tensor
  0.000000   0.000000   0.000000 ~ 566.588135 -53.611449 -26.706546 atm
  0.000000   0.000000   0.000000 ~ -53.611449 409.632800 159.708960 atm
  0.000000   0.000000   0.000000 ~ -26.706546 159.708960 649.350542 atm

 ***** Begining Molecular Dynamics Steps *****
nprs=2313, nprt=7813, nprse=0, nprte=0, <ts/update>=6.1875
<Uer/N>  	=       -5.642468	~       -5.698348 K
<Ke /N>  	=        2.072360	~        2.150451 K
<inter>  	=    -1327.961662	~    -1342.266997 K
<Ham/N>  	=       -3.118552	~       -3.096112 K
<temp>   	=        1.386991	~        1.439256 K
<pres>   	=        2.960737	~        2.816631 K/A^3
<pres>   	=      403.420489	~      383.785139 atm
Pressure Tensor
404.215452  36.027702 -28.484138 ~ 311.831790  83.528560 -75.212325 atm
 36.027702 403.154868  94.363165 ~  83.528560 454.738997  72.848574 atm
-28.484138  94.363165 402.891148 ~ -75.212325  72.848574 384.784630 atm
<dE>     	=        0.019775	~       0.006431
nprs=2352, nprt=7788, nprse=0, nprte=0, <ts/update>=7.37037
<Uer/N>  	=       -5.619628	~       -5.554040 K
<Ke /N>  	=        2.061343	~        2.007331 K
<inter>  	=    -1322.114576	~    -1305.323943 K
<Ham/N>  	=       -3.107316	~       -3.096076 K
<temp>   	=        1.379618	~        1.343469 K
<pres>   	=        3.093726	~        3.456849 K/A^3
<pres>   	=      421.541201	~      471.019163 atm
Pressure Tensor
397.256164  83.981918 -34.269358 ~ 431.919095 139.242240  13.796291 atm
 83.981918 457.736356  65.315005 ~ 139.242240 513.095653 -26.302850 atm
-34.269358  65.315005 409.631083 ~  13.796291 -26.302850 468.042741 atm
<dE>     	=        0.014460	~       0.004693
nprs=2348, nprt=7768, nprse=0, nprte=0, <ts/update>=8.08108
<Uer/N>  	=       -5.597884	~       -5.524687 K
<Ke /N>  	=        2.043902	~        1.977492 K
<inter>  	=    -1316.548036	~    -1297.809776 K
<Ham/N>  	=       -3.103572	~       -3.096051 K
<temp>   	=        1.367945	~        1.323498 K
<pres>   	=        3.229105	~        3.635008 K/A^3
<pres>   	=      439.987520	~      495.294532 atm
Pressure Tensor
436.975116  85.265035 -35.673139 ~ 508.891213  19.430709 -44.758025 atm
 85.265035 476.940582  49.552192 ~  19.430709 569.828448 -28.609598 atm
-35.673139  49.552192 406.046863 ~ -44.758025 -28.609598 407.163933 atm
<dE>     	=        0.011934	~       0.003870
nprs=2362, nprt=7810, nprse=0, nprte=0, <ts/update>=8.67391
<Uer/N>  	=       -5.576514	~       -5.531030 K
<Ke /N>  	=        2.024434	~        1.984393 K
<inter>  	=    -1311.077353	~    -1299.433382 K
<Ham/N>  	=       -3.101697	~       -3.096069 K
<temp>   	=        1.354915	~        1.328117 K
<pres>   	=        3.343994	~        3.636751 K/A^3
<pres>   	=      455.641955	~      495.532092 atm
Pressure Tensor
450.679194  66.219921 -26.177812 ~ 497.968556 -25.725429   4.295552 atm
 66.219921 508.524770  18.195621 ~ -25.725429 540.965118 -106.896367 atm
-26.177812  18.195621 407.721902 ~   4.295552 -106.896367 447.662600 atm
<dE>     	=        0.010390	~       0.003369
nprs=2378, nprt=7832, nprse=0, nprte=0, <ts/update>=9.24074
<Uer/N>  	=       -5.550913	~       -5.426302 K
<Ke /N>  	=        2.000072	~        1.881372 K
<inter>  	=    -1304.523433	~    -1272.623052 K
<Ham/N>  	=       -3.100573	~       -3.096092 K
<temp>   	=        1.338610	~        1.259167 K
<pres>   	=        3.472276	~        4.118336 K/A^3
<pres>   	=      473.121142	~      561.151258 atm
Pressure Tensor
454.092397  48.590166 -26.335648 ~ 519.356129   4.643068 -62.847753 atm
 48.590166 526.824705  -6.831767 ~   4.643068 603.009155 -67.508742 atm
-26.335648  -6.831767 438.446323 ~ -62.847753 -67.508742 561.088490 atm
<dE>     	=        0.009322	~       0.003022
nprs=2360, nprt=7842, nprse=0, nprte=0, <ts/update>=9.66129
<Uer/N>  	=       -5.531373	~       -5.604098 K
<Ke /N>  	=        1.981249	~        2.055733 K
<inter>  	=    -1299.521294	~    -1318.138994 K
<Ham/N>  	=       -3.099822	~       -3.096068 K
<temp>   	=        1.326012	~        1.375863 K
<pres>   	=        3.578903	~        3.386736 K/A^3
<pres>   	=      487.649901	~      461.465770 atm
Pressure Tensor
475.809420  38.848163 -35.221450 ~ 543.756646 -48.399640 -85.287806 atm
 38.848163 537.245573  -6.998125 ~ -48.399640 475.897817  30.741854 atm
-35.221450  -6.998125 449.894708 ~ -85.287806  30.741854 364.742847 atm
<dE>     	=        0.008528	~       0.002764
nprs=2338, nprt=7848, nprse=0, nprte=0, <ts/update>=9.98571
<Uer/N>  	=       -5.537294	~       -5.581590 K
<Ke /N>  	=        1.987441	~        2.034066 K
<inter>  	=    -1301.036958	~    -1312.376929 K
<Ham/N>  	=       -3.099285	~       -3.096060 K
<temp>   	=        1.330157	~        1.361362 K
<pres>   	=        3.550995	~        3.288350 K/A^3
<pres>   	=      483.847146	~      448.060092 atm
Pressure Tensor
480.679905  22.179100 -36.689190 ~ 404.827711 -102.760569  14.513541 atm
 22.179100 522.727568  -8.689557 ~ -102.760569 387.356576 -10.593657 atm
-36.689190  -8.689557 448.133964 ~  14.513541 -10.593657 551.995988 atm
<dE>     	=        0.007907	~       0.002563
nprs=2355, nprt=7843, nprse=0, nprte=0, <ts/update>=10.1139
<Uer/N>  	=       -5.537711	~       -5.526083 K
<Ke /N>  	=        1.988184	~        1.977923 K
<inter>  	=    -1301.143955	~    -1298.167030 K
<Ham/N>  	=       -3.098882	~       -3.096054 K
<temp>   	=        1.330654	~        1.323787 K
<pres>   	=        3.540290	~        3.545772 K/A^3
<pres>   	=      482.388564	~      483.135540 atm
Pressure Tensor
482.936488  12.818863 -37.033332 ~ 580.416253 -50.364288 -25.941293 atm
 12.818863 514.126533  -5.506324 ~ -50.364288 474.963740  28.548248 atm
-37.033332  -5.506324 450.102670 ~ -25.941293  28.548248 394.026627 atm
<dE>     	=        0.007405	~       0.002400
nprs=2359, nprt=7819, nprse=0, nprte=0, <ts/update>=10.3333
<Uer/N>  	=       -5.534233	~       -5.486528 K
<Ke /N>  	=        1.984952	~        1.939149 K
<inter>  	=    -1300.253586	~    -1288.040993 K
<Ham/N>  	=       -3.098568	~       -3.096043 K
<temp>   	=        1.328491	~        1.297835 K
<pres>   	=        3.552856	~        3.743322 K/A^3
<pres>   	=      484.100762	~      510.053067 atm
Pressure Tensor
486.131756   8.513532 -31.893932 ~ 491.504280 -32.084667  29.224498 atm
  8.513532 515.452336   4.046148 ~ -32.084667 553.615709  50.848057 atm
-31.893932   4.046148 450.718194 ~  29.224498  50.848057 485.039213 atm
<dE>     	=        0.006987	~       0.002264
nprs=2378, nprt=7820, nprse=0, nprte=0, <ts/update>=10.5158
<Uer/N>  	=       -5.533310	~       -5.450747 K
<Ke /N>  	=        1.984211	~        1.904161 K
<inter>  	=    -1300.017120	~    -1278.881137 K
<Ham/N>  	=       -3.098317	~       -3.096083 K
<temp>   	=        1.327995	~        1.274419 K
<pres>   	=        3.560311	~        3.921727 K/A^3
<pres>   	=      485.116608	~      534.361964 atm
Pressure Tensor
485.264143   7.752058 -28.957514 ~ 493.556537 -34.273734 -50.815975 atm
  7.752058 516.965788   5.312291 ~ -34.273734 526.208766 -37.019666 atm
-28.957514   5.312291 453.119892 ~ -50.815975 -37.019666 583.320589 atm
<dE>     	=        0.006634	~       0.002149

 *** Finnished Molecular Dynamics Steps ***

Closing files
Saving new initial file (lj.init_N) with final coordinates
:-)
