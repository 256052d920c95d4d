Reading in set from file "lj.set_global"
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
Barostat set up: mvol = 256 and you have 0 barostates
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
# of thermostats      = 1
# of nh chains links  = 3
# of barostats        = 0
Temperature           = 1 K
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
  Pressure = 3.98006 K/A^3 = 542.31 ATM

nprs=2334, nprt=7851, nprse=0, nprte=0, <ts/update>=0
<Uer/N>  	=        0.000000	~       -5.268099 K
<Ke /N>  	=        0.000000	~        1.500000 K
<inter>  	=        0.000000	~    -1232.123279 K
<Ueta/NC>	=        0.000000	~        0.000000 K
<Keta/NC>	=        0.000000	~        0.500000 K
<Ham/N>  	=        0.000000	~       -3.311544 K
<temp>   	=        0.000000	~        1.000000 K
<pres>   	=        0.000000	~        3.980055 K/A^3
<pres>   	=        0.000000	~      542.309574 atm
Pressure Tensor
  0.000000   0.000000   0.000000 ~ 567.019848 -53.584294 -26.679897 atm
  0.000000   0.000000   0.000000 ~ -53.584294 410.098376 159.725105 atm
  0.000000   0.000000   0.000000 ~ -26.679897 159.725105 649.810498 atm

 ***** Begining Molecular Dynamics Steps *****
nprs=2314, nprt=7814, nprse=0, nprte=0, <ts/update>=6.1875
<Uer/N>  	=       -5.641070	~       -5.697210 K
<Ke /N>  	=        2.075353	~        2.149450 K
<inter>  	=    -1327.603650	~    -1341.975558 K
<Ueta/NC>	=        0.123416	~        0.414612 K
<Keta/NC>	=        0.500971	~        0.573642 K
<Ham/N>  	=       -3.106845	~       -3.084394 K
<temp>   	=        1.383569	~        1.432967 K
<pres>   	=        2.968879	~        2.821855 K/A^3
<pres>   	=      404.529938	~      384.496903 atm
Pressure Tensor
405.254213  36.167321 -28.479733 ~ 312.559716  83.703001 -75.363404 atm
 36.167321 404.308246  94.400318 ~  83.703001 455.460377  72.720249 atm
-28.479733  94.400318 404.027356 ~ -75.363404  72.720249 385.470617 atm
<dE>     	=        0.019854	~       0.006453
nprs=2353, nprt=7788, nprse=0, nprte=0, <ts/update>=7.37037
<Uer/N>  	=       -5.619721	~       -5.558871 K
<Ke /N>  	=        2.059013	~        1.991795 K
<inter>  	=    -1322.138361	~    -1306.560761 K
<Ueta/NC>	=        0.579485	~        1.776920 K
<Keta/NC>	=        0.627281	~        0.961671 K
<Ham/N>  	=       -3.095602	~       -3.084351 K
<temp>   	=        1.372676	~        1.327863 K
<pres>   	=        3.092311	~        3.426139 K/A^3
<pres>   	=      421.348415	~      466.834762 atm
Pressure Tensor
396.989111  84.022734 -34.274443 ~ 427.492908 138.789455  13.742019 atm
 84.022734 457.585855  65.226452 ~ 138.789455 508.720102 -26.422058 atm
-34.274443  65.226452 409.470279 ~  13.742019 -26.422058 464.291278 atm
<dE>     	=        0.014518	~       0.004709
nprs=2348, nprt=7770, nprse=0, nprte=0, <ts/update>=8.08108
<Uer/N>  	=       -5.601277	~       -5.542514 K
<Ke /N>  	=        2.033993	~        1.944024 K
<inter>  	=    -1317.416705	~    -1302.373396 K
<Ueta/NC>	=        1.307703	~        3.834960 K
<Keta/NC>	=        0.826528	~        1.526213 K
<Ham/N>  	=       -3.091855	~       -3.084327 K
<temp>   	=        1.355996	~        1.296016 K
<pres>   	=        3.208100	~        3.535295 K/A^3
<pres>   	=      437.125496	~      481.707923 atm
Pressure Tensor
434.027374  85.152373 -35.522161 ~ 496.733823  19.230716 -44.616899 atm
 85.152373 473.896565  49.459064 ~  19.230716 553.730563 -27.852931 atm
-35.522161  49.459064 403.452550 ~ -44.616899 -27.852931 394.659384 atm
<dE>     	=        0.011982	~       0.003884
nprs=2354, nprt=7803, nprse=0, nprte=0, <ts/update>=8.67391
<Uer/N>  	=       -5.584535	~       -5.560354 K
<Ke /N>  	=        2.005139	~        1.925674 K
<inter>  	=    -1313.130747	~    -1306.940471 K
<Ueta/NC>	=        2.256099	~        6.406004 K
<Keta/NC>	=        1.074366	~        2.129888 K
<Ham/N>  	=       -3.089977	~       -3.084338 K
<temp>   	=        1.336759	~        1.283783 K
<pres>   	=        3.296125	~        3.465880 K/A^3
<pres>   	=      449.119423	~      472.249687 atm
Pressure Tensor
444.363164  66.192803 -26.056074 ~ 477.536984 -23.337817   4.604628 atm
 66.192803 501.512651  18.481575 ~ -23.337817 517.394767 -104.463155 atm
-26.056074  18.481575 401.482453 ~   4.604628 -104.463155 421.817309 atm
<dE>     	=        0.010432	~       0.003381
nprs=2378, nprt=7845, nprse=0, nprte=0, <ts/update>=9.24074
<Uer/N>  	=       -5.565252	~       -5.475197 K
<Ke /N>  	=        1.970914	~        1.803354 K
<inter>  	=    -1308.194224	~    -1285.140323 K
<Ueta/NC>	=        3.371128	~        9.238981 K
<Keta/NC>	=        1.339826	~        2.625009 K
<Ham/N>  	=       -3.088851	~       -3.084360 K
<temp>   	=        1.313943	~        1.202236 K
<pres>   	=        3.389176	~        3.841000 K/A^3
<pres>   	=      461.798208	~      523.362403 atm
Pressure Tensor
443.421152  48.722130 -26.089978 ~ 481.475396   4.378730 -60.953859 atm
 48.722130 515.361511  -6.176601 ~   4.378730 568.967867 -67.296663 atm
-26.089978  -6.176601 426.611961 ~ -60.953859 -67.296663 519.643946 atm
<dE>     	=        0.009360	~       0.003033
nprs=2348, nprt=7841, nprse=0, nprte=0, <ts/update>=9.66129
<Uer/N>  	=       -5.551757	~       -5.642389 K
<Ke /N>  	=        1.941047	~        1.930207 K
<inter>  	=    -1304.739558	~    -1327.941307 K
<Ueta/NC>	=        4.585299	~       12.027505 K
<Keta/NC>	=        1.582997	~        3.002259 K
<Ham/N>  	=       -3.088097	~       -3.084331 K
<temp>   	=        1.294031	~        1.286805 K
<pres>   	=        3.460780	~        3.132456 K/A^3
<pres>   	=      471.554854	~      426.818482 atm
Pressure Tensor
459.710446  39.067666 -34.232117 ~ 501.541245 -49.935917 -76.119443 atm
 39.067666 521.098528  -6.702728 ~ -49.935917 446.973343  30.855777 atm
-34.232117  -6.702728 433.855588 ~ -76.119443  30.855777 331.940859 atm
<dE>     	=        0.008563	~       0.002774
nprs=2343, nprt=7833, nprse=0, nprte=0, <ts/update>=9.98571
<Uer/N>  	=       -5.564236	~       -5.632578 K
<Ke /N>  	=        1.936127	~        1.881765 K
<inter>  	=    -1307.934283	~    -1325.429886 K
<Ueta/NC>	=        5.850928	~       14.840968 K
<Keta/NC>	=        1.829289	~        3.561956 K
<Ham/N>  	=       -3.087559	~       -3.084330 K
<temp>   	=        1.290751	~        1.254510 K
<pres>   	=        3.397792	~        2.973217 K/A^3
<pres>   	=      462.972309	~      405.120976 atm
Pressure Tensor
460.005965  22.940684 -35.780774 ~ 365.346484 -96.574829  14.948392 atm
 22.940684 503.400527  -8.155131 ~ -96.574829 350.567256 -22.949422 atm
-35.780774  -8.155131 425.510436 ~  14.948392 -22.949422 499.449188 atm
<dE>     	=        0.007940	~       0.002572
nprs=2347, nprt=7817, nprse=0, nprte=0, <ts/update>=10.2436
<Uer/N>  	=       -5.571120	~       -5.639569 K
<Ke /N>  	=        1.925165	~        1.849573 K
<inter>  	=    -1309.696456	~    -1327.219462 K
<Ueta/NC>	=        7.152159	~       17.640643 K
<Keta/NC>	=        2.075281	~        3.990880 K
<Ham/N>  	=       -3.087155	~       -3.084332 K
<temp>   	=        1.283443	~        1.233049 K
<pres>   	=        3.350969	~        2.966960 K/A^3
<pres>   	=      456.592355	~      404.268516 atm
Pressure Tensor
455.857243  13.677469 -35.877655 ~ 496.818614 -41.084216 -40.549940 atm
 13.677469 490.427896  -5.594401 ~ -41.084216 391.874193  44.931939 atm
-35.877655  -5.594401 423.491925 ~ -40.549940  44.931939 324.112741 atm
<dE>     	=        0.007435	~       0.002408
nprs=2368, nprt=7814, nprse=0, nprte=0, <ts/update>=10.4535
<Uer/N>  	=       -5.573505	~       -5.555947 K
<Ke /N>  	=        1.909604	~        1.732211 K
<inter>  	=    -1310.307182	~    -1305.812300 K
<Ueta/NC>	=        8.470938	~       20.355149 K
<Keta/NC>	=        2.307096	~        4.277574 K
<Ham/N>  	=       -3.086840	~       -3.084312 K
<temp>   	=        1.273069	~        1.154807 K
<pres>   	=        3.327464	~        3.260875 K/A^3
<pres>   	=      453.389587	~      444.316318 atm
Pressure Tensor
455.471438   7.598491 -32.689051 ~ 435.022890 -53.208291  22.657894 atm
  7.598491 485.525886   4.685405 ~ -53.208291 480.179244  37.638697 atm
-32.689051   4.685405 419.171437 ~  22.657894  37.638697 417.746821 atm
<dE>     	=        0.007016	~       0.002272
nprs=2377, nprt=7810, nprse=0, nprte=0, <ts/update>=10.7419
<Uer/N>  	=       -5.577959	~       -5.614812 K
<Ke /N>  	=        1.896314	~        1.756628 K
<inter>  	=    -1311.447396	~    -1320.881568 K
<Ueta/NC>	=        9.791640	~       22.967131 K
<Keta/NC>	=        2.517484	~        4.520710 K
<Ham/N>  	=       -3.086588	~       -3.084340 K
<temp>   	=        1.264210	~        1.171085 K
<pres>   	=        3.297674	~        3.053082 K/A^3
<pres>   	=      449.330496	~      416.003173 atm
Pressure Tensor
454.516392   6.693913 -28.573361 ~ 460.399304  10.351899 -31.449115 atm
  6.693913 480.787389   4.529849 ~  10.351899 422.938688 -40.577239 atm
-28.573361   4.529849 412.687707 ~ -31.449115 -40.577239 364.671526 atm
<dE>     	=        0.006661	~       0.002157

 *** Finnished Molecular Dynamics Steps ***

Closing files
Saving new initial file (lj.init_N) with final coordinates
:-)
