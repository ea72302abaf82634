NAME          config1
ROWS
 N  COST
 E  EQ1
 E  EQ2
 E  EQ3
 E  EQ4
 E  EQ5
 E  EQ6
 E  EQ7
 E  EQ8
 E  EQ9
 E  EQ10
 E  EQ11
 E  EQ12
 E  EQ13
 E  EQ14
 E  EQ15
 E  EQ16
 E  EQ17
 E  EQ18
 E  EQ19
 E  EQ20
 E  EQ21
 E  EQ22
 E  EQ23
 E  EQ24
 E  EQ25
 E  EQ26
 E  EQ27
 E  EQ28
 E  EQ29
 E  EQ30
 E  EQ31
 E  EQ32
 E  EQ33
 E  EQ34
 E  EQ35
 E  EQ36
 E  EQ37
 E  EQ38
 E  EQ39
 E  EQ40
 E  EQ41
 E  EQ42
 E  EQ43
 E  EQ44
 E  EQ45
 E  EQ46
 E  EQ47
 E  EQ48
 E  EQ49
 E  EQ50
 E  EQ51
 E  EQ52
 E  EQ53
 E  EQ54
 E  EQ55
 E  EQ56
 E  EQ57
 E  EQ58
 E  EQ59
 E  EQ60
 E  EQ61
 E  EQ62
 E  EQ63
 E  EQ64
 E  EQ65
 E  EQ66
 E  EQ67
 E  EQ68
 E  EQ69
 E  EQ70
 E  EQ71
 E  EQ72
COLUMNS
    X1        EQ8       1           
    X1        EQ30      1           
    X1        EQ39      -1          
    X1        EQ44      -2          
    X2        EQ12      2           
    X2        EQ31      -2          
    X2        EQ46      -1          
    X2        EQ64      -1          
    X3        EQ40      2           
    X3        EQ69      -2          
    X4        EQ6       2           
    X4        EQ11      2           
    X4        EQ18      -1          
    X4        EQ69      2           
    X5        EQ13      -1          
    X5        EQ50      -1          
    X5        EQ56      -2          
    X6        EQ34      -1          
    X6        EQ42      1           
    X6        EQ48      2           
    X7        EQ27      -2          
    X7        EQ41      -2          
    X7        EQ49      -1          
    X7        EQ51      -2          
    X8        EQ15      2           
    X8        EQ21      -1          
    X8        EQ29      1           
    X9        EQ10      2           
    X9        EQ40      1           
    X9        EQ64      1           
    X9        EQ69      1           
    X10       EQ25      1           
    X10       EQ67      -1          
    X11       EQ6       -1          
    X11       EQ7       2           
    X11       EQ35      -2          
    X11       EQ40      1           
    X12       EQ19      1           
    X12       EQ22      -2          
    X12       EQ55      2           
    X12       EQ69      -2          
    X13       EQ29      2           
    X13       EQ35      1           
    X13       EQ68      2           
    X14       EQ9       2           
    X14       EQ51      -2          
    X14       EQ53      -2          
    X14       EQ68      2           
    X15       EQ28      1           
    X15       EQ41      2           
    X15       EQ71      1           
    X16       EQ38      -1          
    X16       EQ39      1           
    X16       EQ43      2           
    X17       EQ15      2           
    X17       EQ69      1           
    X18       EQ26      -1          
    X18       EQ54      1           
    X18       EQ58      -1          
    X19       EQ28      -2          
    X19       EQ52      -2          
    X20       EQ14      2           
    X20       EQ36      2           
    X21       EQ20      2           
    X21       EQ36      -2          
    X21       EQ54      -1          
    X21       EQ64      -1          
    X22       EQ22      2           
    X22       EQ57      -1          
    X23       EQ17      2           
    X23       EQ19      1           
    X23       EQ62      -2          
    X24       EQ26      -2          
    X24       EQ30      -1          
    X24       EQ44      -1          
    X24       EQ69      -1          
    X25       EQ10      1           
    X25       EQ34      1           
    X25       EQ57      -2          
    X25       EQ68      -1          
    X26       EQ22      -1          
    X26       EQ40      1           
    X26       EQ43      -2          
    X26       EQ55      -1          
    X27       EQ19      2           
    X27       EQ24      1           
    X27       EQ55      -1          
    X27       EQ57      1           
    X28       EQ6       1           
    X28       EQ55      2           
    X29       EQ3       -1          
    X29       EQ40      -2          
    X30       EQ33      -2          
    X30       EQ41      2           
    X30       EQ43      1           
    X30       EQ47      1           
    X31       EQ11      -2          
    X31       EQ50      -2          
    X31       EQ51      -2          
    X31       EQ52      -1          
    X32       EQ13      1           
    X32       EQ54      1           
    X33       EQ23      -1          
    X33       EQ32      1           
    X33       EQ69      2           
    X33       EQ71      1           
    X34       EQ27      1           
    X34       EQ59      -2          
    X34       EQ64      -2          
    X34       EQ67      -1          
    X35       EQ3       2           
    X35       EQ19      -1          
    X35       EQ42      1           
    X35       EQ64      -2          
    X36       EQ5       -2          
    X36       EQ38      1           
    X36       EQ47      1           
    X36       EQ52      -1          
    X37       EQ12      -1          
    X37       EQ13      -2          
    X38       EQ2       2           
    X38       EQ6       -2          
    X38       EQ15      -1          
    X38       EQ26      2           
    X39       EQ10      -2          
    X39       EQ23      2           
    X39       EQ48      2           
    X40       EQ7       2           
    X40       EQ20      2           
    X41       EQ24      -1          
    X41       EQ52      2           
    X41       EQ53      -1          
    X41       EQ71      -1          
    X42       EQ6       -2          
    X42       EQ30      1           
    X42       EQ33      2           
    X42       EQ46      -2          
    X43       EQ8       -2          
    X43       EQ34      -2          
    X43       EQ65      1           
    X43       EQ69      1           
    X44       EQ44      2           
    X44       EQ59      -1          
    X45       EQ3       2           
    X45       EQ37      -1          
    X45       EQ48      -2          
    X45       EQ64      -2          
    X46       EQ17      2           
    X46       EQ29      2           
    X46       EQ47      -1          
    X46       EQ68      -1          
    X47       EQ5       -1          
    X47       EQ12      1           
    X47       EQ46      -1          
    X47       EQ66      2           
    X48       EQ24      2           
    X48       EQ30      1           
    X48       EQ42      -1          
    X48       EQ57      2           
    X49       EQ4       1           
    X49       EQ41      2           
    X50       EQ19      -1          
    X50       EQ29      -2          
    X50       EQ40      2           
    X50       EQ44      1           
    X51       EQ5       1           
    X51       EQ11      2           
    X51       EQ15      -2          
    X51       EQ41      2           
    X52       EQ1       -1          
    X52       EQ8       -2          
    X52       EQ21      -2          
    X52       EQ27      1           
    X53       EQ5       -1          
    X53       EQ11      -1          
    X54       EQ24      -1          
    X54       EQ41      -1          
    X54       EQ58      -1          
    X54       EQ72      2           
    X55       EQ53      2           
    X55       EQ61      1           
    X56       EQ26      -2          
    X56       EQ60      -2          
    X57       EQ14      1           
    X57       EQ30      -1          
    X57       EQ44      -1          
    X57       EQ71      -1          
    X58       EQ16      2           
    X58       EQ33      -1          
    X58       EQ38      -1          
    X58       EQ65      2           
    X59       EQ16      -1          
    X59       EQ22      1           
    X59       EQ36      1           
    X59       EQ50      -1          
    X60       EQ14      2           
    X60       EQ36      1           
    X60       EQ53      -2          
    X61       EQ8       -2          
    X61       EQ60      2           
    X62       EQ13      -2          
    X62       EQ66      -1          
    X63       EQ17      -1          
    X63       EQ24      2           
    X63       EQ31      1           
    X63       EQ32      -1          
    X64       EQ11      2           
    X64       EQ15      -2          
    X64       EQ26      2           
    X64       EQ69      -2          
    X65       EQ4       1           
    X65       EQ38      2           
    X65       EQ49      2           
    X65       EQ60      -2          
    X66       EQ32      1           
    X66       EQ70      -1          
    X67       EQ28      1           
    X67       EQ53      2           
    X68       EQ8       1           
    X68       EQ12      -1          
    X68       EQ18      2           
    X68       EQ45      2           
    X69       EQ22      -1          
    X69       EQ29      -1          
    X69       EQ43      -1          
    X69       EQ64      1           
    X70       EQ2       -1          
    X70       EQ20      2           
    X70       EQ51      1           
    X70       EQ68      -2          
    X71       EQ16      2           
    X71       EQ49      -1          
    X72       EQ13      1           
    X72       EQ29      2           
    X72       EQ38      2           
    X73       EQ12      2           
    X73       EQ28      -2          
    X73       EQ44      2           
    X73       EQ48      2           
    X74       EQ30      1           
    X74       EQ35      -2          
    X74       EQ52      -1          
    X75       EQ30      1           
    X75       EQ31      -1          
    X75       EQ32      -1          
    X75       EQ35      -1          
    X76       EQ19      1           
    X76       EQ52      1           
    X77       EQ12      2           
    X77       EQ20      -2          
    X77       EQ56      -1          
    X77       EQ71      -1          
    X78       EQ3       -2          
    X78       EQ40      -2          
    X78       EQ45      -2          
    X78       EQ58      -1          
    X79       EQ10      1           
    X79       EQ32      -2          
    X79       EQ41      -2          
    X79       EQ60      -1          
    X80       EQ31      -2          
    X80       EQ39      2           
    X80       EQ46      -2          
    X80       EQ72      -2          
    X81       EQ8       1           
    X81       EQ12      -1          
    X81       EQ29      -1          
    X81       EQ65      -1          
    X82       EQ26      -2          
    X82       EQ40      1           
    X83       EQ12      2           
    X83       EQ39      -2          
    X83       EQ52      -1          
    X83       EQ57      -1          
    X84       EQ19      1           
    X84       EQ67      2           
    X85       EQ6       2           
    X85       EQ43      -1          
    X85       EQ52      -1          
    X85       EQ54      -1          
    X86       EQ25      2           
    X86       EQ45      1           
    X87       EQ31      -1          
    X87       EQ44      -1          
    X87       EQ64      -1          
    X87       EQ66      1           
    X88       EQ64      -1          
    X88       EQ70      -2          
    X89       EQ2       -1          
    X89       EQ19      -2          
    X89       EQ41      2           
    X90       EQ12      1           
    X90       EQ41      -1          
    X90       EQ45      1           
    X91       EQ15      2           
    X91       EQ29      1           
    X91       EQ41      -2          
    X91       EQ62      2           
    X92       EQ63      2           
    X92       EQ72      -2          
    X93       EQ48      -2          
    X93       EQ62      -2          
    X93       EQ68      -2          
    X94       EQ22      2           
    X94       EQ54      -2          
    X95       EQ23      1           
    X95       EQ67      1           
RHS
BOUNDS
 UP BOUND     X1        30.6412     
 LO BOUND     X2        -10.2909    
 UP BOUND     X2        21.6416     
 FX BOUND     X3        5.5277      
 LO BOUND     X4        -50.0776    
 UP BOUND     X4        38.6451     
 LO BOUND     X5        -44.886     
 UP BOUND     X5        15.4884     
 LO BOUND     X6        -39.8004    
 UP BOUND     X6        24.7475     
 FX BOUND     X7        0.5089      
 LO BOUND     X8        -37.3823    
 UP BOUND     X8        28.0957     
 UP BOUND     X9        40.6729     
 LO BOUND     X10       -35.9514    
 UP BOUND     X10       54.4626     
 LO BOUND     X11       -28.4484    
 UP BOUND     X11       25.5446     
 LO BOUND     X12       -17.0076    
 UP BOUND     X12       16.5068     
 FX BOUND     X13       0.7677      
 LO BOUND     X14       -45.8715    
 UP BOUND     X14       21.7288     
 UP BOUND     X15       10.9098     
 LO BOUND     X16       -13.1747    
 UP BOUND     X16       28.4036     
 LO BOUND     X17       -40.3108    
 UP BOUND     X17       42.6609     
 LO BOUND     X18       -20.0568    
 UP BOUND     X18       43.5596     
 LO BOUND     X19       -38.817     
 UP BOUND     X19       43.8025     
 LO BOUND     X20       -21.672     
 UP BOUND     X20       20.9025     
 UP BOUND     X21       17.2474     
 LO BOUND     X22       -15.0817    
 UP BOUND     X22       7.8076      
 LO BOUND     X23       -11.8883    
 UP BOUND     X23       20.9349     
 LO BOUND     X24       -8.8196     
 UP BOUND     X24       40.8463     
 LO BOUND     X25       -10.2721    
 UP BOUND     X25       13.396      
 UP BOUND     X26       44.2195     
 LO BOUND     X27       -42.4682    
 UP BOUND     X27       37.0913     
 LO BOUND     X28       -25.9349    
 UP BOUND     X28       16.4089     
 UP BOUND     X29       39.4565     
 LO BOUND     X30       -28.0814    
 UP BOUND     X30       42.4296     
 LO BOUND     X31       -31.5207    
 UP BOUND     X31       44.8768     
 LO BOUND     X32       -28.5518    
 UP BOUND     X32       11.0945     
 LO BOUND     X33       -14.4644    
 UP BOUND     X33       41.8182     
 LO BOUND     X34       -23.5301    
 UP BOUND     X34       40.931      
 LO BOUND     X35       -43.4398    
 UP BOUND     X35       16.7162     
 LO BOUND     X36       -17.2784    
 UP BOUND     X36       23.2458     
 LO BOUND     X37       -23.4177    
 UP BOUND     X37       52.3304     
 LO BOUND     X38       -51.1617    
 UP BOUND     X38       16.9899     
 LO BOUND     X39       -15.8912    
 UP BOUND     X39       17.5888     
 LO BOUND     X40       -36.7862    
 UP BOUND     X40       38.2621     
 LO BOUND     X41       -26.9243    
 UP BOUND     X41       10.8293     
 UP BOUND     X42       27.134      
 LO BOUND     X43       -20.9528    
 UP BOUND     X43       21.3485     
 LO BOUND     X44       -31.4417    
 UP BOUND     X44       27.4438     
 LO BOUND     X45       -21.5729    
 UP BOUND     X45       32.1463     
 LO BOUND     X46       -38.4436    
 UP BOUND     X46       39.3379     
 LO BOUND     X47       -51.1848    
 UP BOUND     X47       19.44       
 LO BOUND     X48       -35.3112    
 UP BOUND     X48       43.0327     
 LO BOUND     X49       -50.4282    
 UP BOUND     X49       41.3231     
 LO BOUND     X50       -36.5459    
 UP BOUND     X50       48.3977     
 LO BOUND     X51       -34.3346    
 UP BOUND     X51       33.1002     
 LO BOUND     X52       -45.2624    
 UP BOUND     X52       37.8539     
 LO BOUND     X53       -14.1627    
 UP BOUND     X53       27.4544     
 LO BOUND     X54       -45.5943    
 UP BOUND     X54       21.3515     
 LO BOUND     X55       -22.5686    
 UP BOUND     X55       34.3349     
 LO BOUND     X56       -29.5165    
 UP BOUND     X56       45.4274     
 LO BOUND     X57       -11.088     
 UP BOUND     X57       43.2204     
 LO BOUND     X58       -42.2706    
 UP BOUND     X58       17.53       
 LO BOUND     X59       -16.6571    
 UP BOUND     X59       18.7681     
 LO BOUND     X60       -36.3886    
 UP BOUND     X60       39.3559     
 LO BOUND     X61       -15.0969    
 UP BOUND     X61       36.4599     
 LO BOUND     X62       -28.0406    
 UP BOUND     X62       36.6328     
 UP BOUND     X63       15.2614     
 LO BOUND     X64       -21.2582    
 UP BOUND     X64       38.1655     
 FX BOUND     X65       1.2823      
 LO BOUND     X66       -33.4463    
 UP BOUND     X66       23.5137     
 LO BOUND     X67       -48.2098    
 UP BOUND     X67       20.0021     
 FX BOUND     X68       -1.4951     
 LO BOUND     X69       -11.6184    
 UP BOUND     X69       48.7224     
 LO BOUND     X70       -23.8292    
 UP BOUND     X70       5.5816      
 LO BOUND     X71       -6.6688     
 UP BOUND     X71       21.2392     
 LO BOUND     X72       -38.2665    
 UP BOUND     X72       33.3525     
 LO BOUND     X73       -16.272     
 UP BOUND     X73       11.9025     
 UP BOUND     X74       7.8568      
 LO BOUND     X75       -45.7528    
 UP BOUND     X75       7.0305      
 LO BOUND     X76       -49.489     
 UP BOUND     X76       38.3621     
 LO BOUND     X77       -47.8784    
 UP BOUND     X77       18.615      
 LO BOUND     X78       -49.3427    
 UP BOUND     X78       17.0782     
 UP BOUND     X79       41.878      
 LO BOUND     X80       -23.7973    
 UP BOUND     X80       23.1417     
 LO BOUND     X81       -7.0682     
 UP BOUND     X81       37.8538     
 LO BOUND     X82       -16.1382    
 UP BOUND     X82       15.0919     
 UP BOUND     X83       21.1982     
 LO BOUND     X84       -41.0767    
 UP BOUND     X84       24.8689     
 LO BOUND     X85       -27.0018    
 UP BOUND     X85       10.0284     
 LO BOUND     X86       -7.4316     
 UP BOUND     X86       6.1118      
 UP BOUND     X87       47.0108     
 LO BOUND     X88       -42.3052    
 UP BOUND     X88       40.9707     
 LO BOUND     X89       -17.5434    
 UP BOUND     X89       20.3224     
 UP BOUND     X90       36.7008     
 LO BOUND     X91       -19.275     
 UP BOUND     X91       42.3934     
 LO BOUND     X92       -32.4178    
 UP BOUND     X92       22.5826     
 UP BOUND     X93       39.3011     
 LO BOUND     X94       -15.418     
 UP BOUND     X94       38.3978     
 LO BOUND     X95       -31.3182    
 UP BOUND     X95       9.1712      
ENDATA
