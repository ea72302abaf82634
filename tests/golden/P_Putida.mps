NAME          Putida                                         
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
 E  EQ73                                                     
 E  EQ74                                                     
 E  EQ75                                                     
 E  EQ76                                                     
 E  EQ77                                                     
 E  EQ78                                                     
 E  EQ79                                                     
 E  EQ80                                                     
 E  EQ81                                                     
 E  EQ82                                                     
 E  EQ83                                                     
 E  EQ84                                                     
 E  EQ85                                                     
 E  EQ86                                                     
 E  EQ87                                                     
 E  EQ88                                                     
 E  EQ89                                                     
 E  EQ90                                                     
 E  EQ91                                                     
 E  EQ92                                                     
 E  EQ93                                                     
 E  EQ94                                                     
 E  EQ95                                                     
 E  EQ96                                                     
 E  EQ97                                                     
 E  EQ98                                                     
 E  EQ99                                                     
 E  EQ100                                                    
 E  EQ101                                                    
 E  EQ102                                                    
 E  EQ103                                                    
 E  EQ104                                                    
 E  EQ105                                                    
 E  EQ106                                                    
 E  EQ107                                                    
 E  EQ108                                                    
 E  EQ109                                                    
 E  EQ110                                                    
 E  EQ111                                                    
 E  EQ112                                                    
 E  EQ113                                                    
 E  EQ114                                                    
 E  EQ115                                                    
 E  EQ116                                                    
 E  EQ117                                                    
 E  EQ118                                                    
 E  EQ119                                                    
 E  EQ120                                                    
 E  EQ121                                                    
 E  EQ122                                                    
 E  EQ123                                                    
 E  EQ124                                                    
 E  EQ125                                                    
 E  EQ126                                                    
 E  EQ127                                                    
 E  EQ128                                                    
 E  EQ129                                                    
 E  EQ130                                                    
 E  EQ131                                                    
 E  EQ132                                                    
 E  EQ133                                                    
 E  EQ134                                                    
 E  EQ135                                                    
 E  EQ136                                                    
 E  EQ137                                                    
 E  EQ138                                                    
 E  EQ139                                                    
 E  EQ140                                                    
 E  EQ141                                                    
 E  EQ142                                                    
 E  EQ143                                                    
 E  EQ144                                                    
 E  EQ145                                                    
 E  EQ146                                                    
 E  EQ147                                                    
 E  EQ148                                                    
 E  EQ149                                                    
 E  EQ150                                                    
 E  EQ151                                                    
 E  EQ152                                                    
 E  EQ153                                                    
 E  EQ154                                                    
 E  EQ155                                                    
 E  EQ156                                                    
 E  EQ157                                                    
 E  EQ158                                                    
 E  EQ159                                                    
 E  EQ160                                                    
 E  EQ161                                                    
 E  EQ162                                                    
 E  EQ163                                                    
 E  EQ164                                                    
 E  EQ165                                                    
 E  EQ166                                                    
 E  EQ167                                                    
 E  EQ168                                                    
 E  EQ169                                                    
 E  EQ170                                                    
 E  EQ171                                                    
 E  EQ172                                                    
 E  EQ173                                                    
 E  EQ174                                                    
 E  EQ175                                                    
 E  EQ176                                                    
 E  EQ177                                                    
 E  EQ178                                                    
 E  EQ179                                                    
 E  EQ180                                                    
 E  EQ181                                                    
 E  EQ182                                                    
 E  EQ183                                                    
 E  EQ184                                                    
 E  EQ185                                                    
 E  EQ186                                                    
 E  EQ187                                                    
 E  EQ188                                                    
 E  EQ189                                                    
 E  EQ190                                                    
 E  EQ191                                                    
 E  EQ192                                                    
 E  EQ193                                                    
 E  EQ194                                                    
 E  EQ195                                                    
 E  EQ196                                                    
 E  EQ197                                                    
 E  EQ198                                                    
 E  EQ199                                                    
 E  EQ200                                                    
 E  EQ201                                                    
 E  EQ202                                                    
 E  EQ203                                                    
 E  EQ204                                                    
 E  EQ205                                                    
 E  EQ206                                                    
 E  EQ207                                                    
 E  EQ208                                                    
 E  EQ209                                                    
 E  EQ210                                                    
 E  EQ211                                                    
 E  EQ212                                                    
 E  EQ213                                                    
 E  EQ214                                                    
 E  EQ215                                                    
 E  EQ216                                                    
 E  EQ217                                                    
 E  EQ218                                                    
 E  EQ219                                                    
 E  EQ220                                                    
 E  EQ221                                                    
 E  EQ222                                                    
 E  EQ223                                                    
 E  EQ224                                                    
 E  EQ225                                                    
 E  EQ226                                                    
 E  EQ227                                                    
 E  EQ228                                                    
 E  EQ229                                                    
 E  EQ230                                                    
 E  EQ231                                                    
 E  EQ232                                                    
 E  EQ233                                                    
 E  EQ234                                                    
 E  EQ235                                                    
 E  EQ236                                                    
 E  EQ237                                                    
 E  EQ238                                                    
 E  EQ239                                                    
 E  EQ240                                                    
 E  EQ241                                                    
 E  EQ242                                                    
 E  EQ243                                                    
 E  EQ244                                                    
 E  EQ245                                                    
 E  EQ246                                                    
 E  EQ247                                                    
 E  EQ248                                                    
 E  EQ249                                                    
 E  EQ250                                                    
 E  EQ251                                                    
 E  EQ252                                                    
 E  EQ253                                                    
 E  EQ254                                                    
 E  EQ255                                                    
 E  EQ256                                                    
 E  EQ257                                                    
 E  EQ258                                                    
 E  EQ259                                                    
 E  EQ260                                                    
 E  EQ261                                                    
 E  EQ262                                                    
 E  EQ263                                                    
 E  EQ264                                                    
 E  EQ265                                                    
 E  EQ266                                                    
 E  EQ267                                                    
 E  EQ268                                                    
 E  EQ269                                                    
 E  EQ270                                                    
 E  EQ271                                                    
 E  EQ272                                                    
 E  EQ273                                                    
 E  EQ274                                                    
 E  EQ275                                                    
 E  EQ276                                                    
 E  EQ277                                                    
 E  EQ278                                                    
 E  EQ279                                                    
 E  EQ280                                                    
 E  EQ281                                                    
 E  EQ282                                                    
 E  EQ283                                                    
 E  EQ284                                                    
 E  EQ285                                                    
 E  EQ286                                                    
 E  EQ287                                                    
 E  EQ288                                                    
 E  EQ289                                                    
 E  EQ290                                                    
 E  EQ291                                                    
 E  EQ292                                                    
 E  EQ293                                                    
 E  EQ294                                                    
 E  EQ295                                                    
 E  EQ296                                                    
 E  EQ297                                                    
 E  EQ298                                                    
 E  EQ299                                                    
 E  EQ300                                                    
 E  EQ301                                                    
 E  EQ302                                                    
 E  EQ303                                                    
 E  EQ304                                                    
 E  EQ305                                                    
 E  EQ306                                                    
 E  EQ307                                                    
 E  EQ308                                                    
 E  EQ309                                                    
 E  EQ310                                                    
 E  EQ311                                                    
 E  EQ312                                                    
 E  EQ313                                                    
 E  EQ314                                                    
 E  EQ315                                                    
 E  EQ316                                                    
 E  EQ317                                                    
 E  EQ318                                                    
 E  EQ319                                                    
 E  EQ320                                                    
 E  EQ321                                                    
 E  EQ322                                                    
 E  EQ323                                                    
 E  EQ324                                                    
 E  EQ325                                                    
 E  EQ326                                                    
 E  EQ327                                                    
 E  EQ328                                                    
 E  EQ329                                                    
 E  EQ330                                                    
 E  EQ331                                                    
 E  EQ332                                                    
 E  EQ333                                                    
 E  EQ334                                                    
 E  EQ335                                                    
 E  EQ336                                                    
 E  EQ337                                                    
 E  EQ338                                                    
 E  EQ339                                                    
 E  EQ340                                                    
 E  EQ341                                                    
 E  EQ342                                                    
 E  EQ343                                                    
 E  EQ344                                                    
 E  EQ345                                                    
 E  EQ346                                                    
 E  EQ347                                                    
 E  EQ348                                                    
 E  EQ349                                                    
 E  EQ350                                                    
 E  EQ351                                                    
 E  EQ352                                                    
 E  EQ353                                                    
 E  EQ354                                                    
 E  EQ355                                                    
 E  EQ356                                                    
 E  EQ357                                                    
 E  EQ358                                                    
 E  EQ359                                                    
 E  EQ360                                                    
 E  EQ361                                                    
 E  EQ362                                                    
 E  EQ363                                                    
 E  EQ364                                                    
 E  EQ365                                                    
 E  EQ366                                                    
 E  EQ367                                                    
 E  EQ368                                                    
 E  EQ369                                                    
 E  EQ370                                                    
 E  EQ371                                                    
 E  EQ372                                                    
 E  EQ373                                                    
 E  EQ374                                                    
 E  EQ375                                                    
 E  EQ376                                                    
 E  EQ377                                                    
 E  EQ378                                                    
 E  EQ379                                                    
 E  EQ380                                                    
 E  EQ381                                                    
 E  EQ382                                                    
 E  EQ383                                                    
 E  EQ384                                                    
 E  EQ385                                                    
 E  EQ386                                                    
 E  EQ387                                                    
 E  EQ388                                                    
 E  EQ389                                                    
 E  EQ390                                                    
 E  EQ391                                                    
 E  EQ392                                                    
 E  EQ393                                                    
 E  EQ394                                                    
 E  EQ395                                                    
 E  EQ396                                                    
 E  EQ397                                                    
 E  EQ398                                                    
 E  EQ399                                                    
 E  EQ400                                                    
 E  EQ401                                                    
 E  EQ402                                                    
 E  EQ403                                                    
 E  EQ404                                                    
 E  EQ405                                                    
 E  EQ406                                                    
 E  EQ407                                                    
 E  EQ408                                                    
 E  EQ409                                                    
 E  EQ410                                                    
 E  EQ411                                                    
 E  EQ412                                                    
 E  EQ413                                                    
 E  EQ414                                                    
 E  EQ415                                                    
 E  EQ416                                                    
 E  EQ417                                                    
 E  EQ418                                                    
 E  EQ419                                                    
 E  EQ420                                                    
 E  EQ421                                                    
 E  EQ422                                                    
 E  EQ423                                                    
 E  EQ424                                                    
 E  EQ425                                                    
 E  EQ426                                                    
 E  EQ427                                                    
 E  EQ428                                                    
 E  EQ429                                                    
 E  EQ430                                                    
 E  EQ431                                                    
 E  EQ432                                                    
 E  EQ433                                                    
 E  EQ434                                                    
 E  EQ435                                                    
 E  EQ436                                                    
 E  EQ437                                                    
 E  EQ438                                                    
 E  EQ439                                                    
 E  EQ440                                                    
 E  EQ441                                                    
 E  EQ442                                                    
 E  EQ443                                                    
 E  EQ444                                                    
 E  EQ445                                                    
 E  EQ446                                                    
 E  EQ447                                                    
 E  EQ448                                                    
 E  EQ449                                                    
 E  EQ450                                                    
 E  EQ451                                                    
 E  EQ452                                                    
 E  EQ453                                                    
 E  EQ454                                                    
 E  EQ455                                                    
 E  EQ456                                                    
 E  EQ457                                                    
 E  EQ458                                                    
 E  EQ459                                                    
 E  EQ460                                                    
 E  EQ461                                                    
 E  EQ462                                                    
 E  EQ463                                                    
 E  EQ464                                                    
 E  EQ465                                                    
 E  EQ466                                                    
 E  EQ467                                                    
 E  EQ468                                                    
 E  EQ469                                                    
 E  EQ470                                                    
 E  EQ471                                                    
 E  EQ472                                                    
 E  EQ473                                                    
 E  EQ474                                                    
 E  EQ475                                                    
 E  EQ476                                                    
 E  EQ477                                                    
 E  EQ478                                                    
 E  EQ479                                                    
 E  EQ480                                                    
 E  EQ481                                                    
 E  EQ482                                                    
 E  EQ483                                                    
 E  EQ484                                                    
 E  EQ485                                                    
 E  EQ486                                                    
 E  EQ487                                                    
 E  EQ488                                                    
 E  EQ489                                                    
 E  EQ490                                                    
 E  EQ491                                                    
 E  EQ492                                                    
 E  EQ493                                                    
 E  EQ494                                                    
 E  EQ495                                                    
 E  EQ496                                                    
 E  EQ497                                                    
 E  EQ498                                                    
 E  EQ499                                                    
 E  EQ500                                                    
 E  EQ501                                                    
 E  EQ502                                                    
 E  EQ503                                                    
 E  EQ504                                                    
 E  EQ505                                                    
 E  EQ506                                                    
 E  EQ507                                                    
 E  EQ508                                                    
 E  EQ509                                                    
 E  EQ510                                                    
 E  EQ511                                                    
 E  EQ512                                                    
 E  EQ513                                                    
 E  EQ514                                                    
 E  EQ515                                                    
 E  EQ516                                                    
 E  EQ517                                                    
 E  EQ518                                                    
 E  EQ519                                                    
 E  EQ520                                                    
 E  EQ521                                                    
 E  EQ522                                                    
 E  EQ523                                                    
 E  EQ524                                                    
 E  EQ525                                                    
 E  EQ526                                                    
 E  EQ527                                                    
 E  EQ528                                                    
 E  EQ529                                                    
 E  EQ530                                                    
 E  EQ531                                                    
 E  EQ532                                                    
 E  EQ533                                                    
 E  EQ534                                                    
 E  EQ535                                                    
 E  EQ536                                                    
 E  EQ537                                                    
 E  EQ538                                                    
 E  EQ539                                                    
 E  EQ540                                                    
 E  EQ541                                                    
 E  EQ542                                                    
 E  EQ543                                                    
 E  EQ544                                                    
 E  EQ545                                                    
 E  EQ546                                                    
 E  EQ547                                                    
 E  EQ548                                                    
 E  EQ549                                                    
 E  EQ550                                                    
 E  EQ551                                                    
 E  EQ552                                                    
 E  EQ553                                                    
 E  EQ554                                                    
 E  EQ555                                                    
 E  EQ556                                                    
 E  EQ557                                                    
 E  EQ558                                                    
 E  EQ559                                                    
 E  EQ560                                                    
 E  EQ561                                                    
 E  EQ562                                                    
 E  EQ563                                                    
 E  EQ564                                                    
 E  EQ565                                                    
 E  EQ566                                                    
 E  EQ567                                                    
 E  EQ568                                                    
 E  EQ569                                                    
 E  EQ570                                                    
 E  EQ571                                                    
 E  EQ572                                                    
 E  EQ573                                                    
 E  EQ574                                                    
 E  EQ575                                                    
 E  EQ576                                                    
 E  EQ577                                                    
 E  EQ578                                                    
 E  EQ579                                                    
 E  EQ580                                                    
 E  EQ581                                                    
 E  EQ582                                                    
 E  EQ583                                                    
 E  EQ584                                                    
 E  EQ585                                                    
 E  EQ586                                                    
 E  EQ587                                                    
 E  EQ588                                                    
 E  EQ589                                                    
 E  EQ590                                                    
 E  EQ591                                                    
 E  EQ592                                                    
 E  EQ593                                                    
 E  EQ594                                                    
 E  EQ595                                                    
 E  EQ596                                                    
 E  EQ597                                                    
 E  EQ598                                                    
 E  EQ599                                                    
 E  EQ600                                                    
 E  EQ601                                                    
 E  EQ602                                                    
 E  EQ603                                                    
 E  EQ604                                                    
 E  EQ605                                                    
 E  EQ606                                                    
 E  EQ607                                                    
 E  EQ608                                                    
 E  EQ609                                                    
 E  EQ610                                                    
 E  EQ611                                                    
 E  EQ612                                                    
 E  EQ613                                                    
 E  EQ614                                                    
 E  EQ615                                                    
 E  EQ616                                                    
 E  EQ617                                                    
 E  EQ618                                                    
 E  EQ619                                                    
 E  EQ620                                                    
 E  EQ621                                                    
 E  EQ622                                                    
 E  EQ623                                                    
 E  EQ624                                                    
 E  EQ625                                                    
 E  EQ626                                                    
 E  EQ627                                                    
 E  EQ628                                                    
 E  EQ629                                                    
 E  EQ630                                                    
 E  EQ631                                                    
 E  EQ632                                                    
 E  EQ633                                                    
 E  EQ634                                                    
 E  EQ635                                                    
 E  EQ636                                                    
 E  EQ637                                                    
 E  EQ638                                                    
 E  EQ639                                                    
 E  EQ640                                                    
 E  EQ641                                                    
 E  EQ642                                                    
 E  EQ643                                                    
 E  EQ644                                                    
 E  EQ645                                                    
 E  EQ646                                                    
 E  EQ647                                                    
 E  EQ648                                                    
 E  EQ649                                                    
 E  EQ650                                                    
 E  EQ651                                                    
 E  EQ652                                                    
 E  EQ653                                                    
 E  EQ654                                                    
 E  EQ655                                                    
 E  EQ656                                                    
 E  EQ657                                                    
 E  EQ658                                                    
 E  EQ659                                                    
 E  EQ660                                                    
 E  EQ661                                                    
 E  EQ662                                                    
 E  EQ663                                                    
 E  EQ664                                                    
 E  EQ665                                                    
 E  EQ666                                                    
 E  EQ667                                                    
 E  EQ668                                                    
 E  EQ669                                                    
 E  EQ670                                                    
 E  EQ671                                                    
 E  EQ672                                                    
 E  EQ673                                                    
 E  EQ674                                                    
 E  EQ675                                                    
 E  EQ676                                                    
 E  EQ677                                                    
 E  EQ678                                                    
 E  EQ679                                                    
 E  EQ680                                                    
 E  EQ681                                                    
 E  EQ682                                                    
 E  EQ683                                                    
 E  EQ684                                                    
 E  EQ685                                                    
 E  EQ686                                                    
 E  EQ687                                                    
 E  EQ688                                                    
 E  EQ689                                                    
 E  EQ690                                                    
 E  EQ691                                                    
 E  EQ692                                                    
 E  EQ693                                                    
 E  EQ694                                                    
 E  EQ695                                                    
 E  EQ696                                                    
 E  EQ697                                                    
 E  EQ698                                                    
 E  EQ699                                                    
 E  EQ700                                                    
 E  EQ701                                                    
 E  EQ702                                                    
 E  EQ703                                                    
 E  EQ704                                                    
 E  EQ705                                                    
 E  EQ706                                                    
 E  EQ707                                                    
 E  EQ708                                                    
 E  EQ709                                                    
 E  EQ710                                                    
 E  EQ711                                                    
 E  EQ712                                                    
 E  EQ713                                                    
 E  EQ714                                                    
 E  EQ715                                                    
 E  EQ716                                                    
 E  EQ717                                                    
 E  EQ718                                                    
 E  EQ719                                                    
 E  EQ720                                                    
 E  EQ721                                                    
 E  EQ722                                                    
 E  EQ723                                                    
 E  EQ724                                                    
 E  EQ725                                                    
 E  EQ726                                                    
 E  EQ727                                                    
 E  EQ728                                                    
 E  EQ729                                                    
 E  EQ730                                                    
 E  EQ731                                                    
 E  EQ732                                                    
 E  EQ733                                                    
 E  EQ734                                                    
 E  EQ735                                                    
 E  EQ736                                                    
 E  EQ737                                                    
 E  EQ738                                                    
 E  EQ739                                                    
 E  EQ740                                                    
 E  EQ741                                                    
 E  EQ742                                                    
 E  EQ743                                                    
 E  EQ744                                                    
 E  EQ745                                                    
 E  EQ746                                                    
 E  EQ747                                                    
 E  EQ748                                                    
 E  EQ749                                                    
 E  EQ750                                                    
 E  EQ751                                                    
 E  EQ752                                                    
 E  EQ753                                                    
 E  EQ754                                                    
 E  EQ755                                                    
 E  EQ756                                                    
 E  EQ757                                                    
 E  EQ758                                                    
 E  EQ759                                                    
 E  EQ760                                                    
 E  EQ761                                                    
 E  EQ762                                                    
 E  EQ763                                                    
 E  EQ764                                                    
 E  EQ765                                                    
 E  EQ766                                                    
 E  EQ767                                                    
 E  EQ768                                                    
 E  EQ769                                                    
 E  EQ770                                                    
 E  EQ771                                                    
 E  EQ772                                                    
 E  EQ773                                                    
 E  EQ774                                                    
 E  EQ775                                                    
 E  EQ776                                                    
 E  EQ777                                                    
 E  EQ778                                                    
 E  EQ779                                                    
 E  EQ780                                                    
 E  EQ781                                                    
 E  EQ782                                                    
 E  EQ783                                                    
 E  EQ784                                                    
 E  EQ785                                                    
 E  EQ786                                                    
 E  EQ787                                                    
 E  EQ788                                                    
 E  EQ789                                                    
 E  EQ790                                                    
 E  EQ791                                                    
 E  EQ792                                                    
 E  EQ793                                                    
 E  EQ794                                                    
 E  EQ795                                                    
 E  EQ796                                                    
 E  EQ797                                                    
 E  EQ798                                                    
 E  EQ799                                                    
 E  EQ800                                                    
 E  EQ801                                                    
 E  EQ802                                                    
 E  EQ803                                                    
 E  EQ804                                                    
 E  EQ805                                                    
 E  EQ806                                                    
 E  EQ807                                                    
 E  EQ808                                                    
 E  EQ809                                                    
 E  EQ810                                                    
 E  EQ811                                                    
 E  EQ812                                                    
 E  EQ813                                                    
 E  EQ814                                                    
 E  EQ815                                                    
 E  EQ816                                                    
 E  EQ817                                                    
 E  EQ818                                                    
 E  EQ819                                                    
 E  EQ820                                                    
 E  EQ821                                                    
 E  EQ822                                                    
 E  EQ823                                                    
 E  EQ824                                                    
 E  EQ825                                                    
 E  EQ826                                                    
 E  EQ827                                                    
 E  EQ828                                                    
 E  EQ829                                                    
 E  EQ830                                                    
 E  EQ831                                                    
 E  EQ832                                                    
 E  EQ833                                                    
 E  EQ834                                                    
 E  EQ835                                                    
 E  EQ836                                                    
 E  EQ837                                                    
 E  EQ838                                                    
 E  EQ839                                                    
 E  EQ840                                                    
 E  EQ841                                                    
 E  EQ842                                                    
 E  EQ843                                                    
 E  EQ844                                                    
 E  EQ845                                                    
 E  EQ846                                                    
 E  EQ847                                                    
 E  EQ848                                                    
 E  EQ849                                                    
 E  EQ850                                                    
 E  EQ851                                                    
 E  EQ852                                                    
 E  EQ853                                                    
 E  EQ854                                                    
 E  EQ855                                                    
 E  EQ856                                                    
 E  EQ857                                                    
 E  EQ858                                                    
 E  EQ859                                                    
 E  EQ860                                                    
 E  EQ861                                                    
 E  EQ862                                                    
 E  EQ863                                                    
 E  EQ864                                                    
 E  EQ865                                                    
 E  EQ866                                                    
 E  EQ867                                                    
 E  EQ868                                                    
 E  EQ869                                                    
 E  EQ870                                                    
 E  EQ871                                                    
 E  EQ872                                                    
 E  EQ873                                                    
 E  EQ874                                                    
 E  EQ875                                                    
 E  EQ876                                                    
 E  EQ877                                                    
 E  EQ878                                                    
 E  EQ879                                                    
 E  EQ880                                                    
 E  EQ881                                                    
 E  EQ882                                                    
 E  EQ883                                                    
 E  EQ884                                                    
 E  EQ885                                                    
 E  EQ886                                                    
 E  EQ887                                                    
 E  EQ888                                                    
 E  EQ889                                                    
 E  EQ890                                                    
 E  EQ891                                                    
 E  EQ892                                                    
 E  EQ893                                                    
 E  EQ894                                                    
 E  EQ895                                                    
 E  EQ896                                                    
 E  EQ897                                                    
 E  EQ898                                                    
 E  EQ899                                                    
 E  EQ900                                                    
 E  EQ901                                                    
 E  EQ902                                                    
 E  EQ903                                                    
 E  EQ904                                                    
 E  EQ905                                                    
 E  EQ906                                                    
 E  EQ907                                                    
 E  EQ908                                                    
 E  EQ909                                                    
 E  EQ910                                                    
 E  EQ911                                                    
COLUMNS                                                      
    X1        COST      1              EQ144     -0.05       
    X1        EQ182     -5E-05         EQ201     45.5608     
    X1        EQ211     -0.488         EQ217     -0.001      
    X1        EQ223     -0.281         EQ226     -0.229      
    X1        EQ227     -0.229         EQ231     -45.7318    
    X1        EQ269     -6E-06         EQ273     -0.0005     
    X1        EQ274     -0.0005        EQ275     -0.0005     
    X1        EQ276     -0.0005        EQ281     -0.126      
    X1        EQ282     -0.087         EQ291     -0.0247     
    X1        EQ299     -0.0254        EQ307     -0.0254     
    X1        EQ331     -0.0247        EQ342     -1E-05      
    X1        EQ375     -0.25          EQ378     -0.25       
    X1        EQ388     -0.582         EQ401     -0.203      
    X1        EQ403     45.56035       EQ407     -45.5608    
    X1        EQ414     -0.0005        EQ419     -0.09       
    X1        EQ435     -0.276         EQ446     -0.428      
    X1        EQ450     -0.326         EQ461     -0.146      
    X1        EQ472     -0.00215       EQ473     -5E-05      
    X1        EQ474     -0.00013       EQ475     -0.0004     
    X1        EQ520     -0.028         EQ521     -0.0005     
    X1        EQ522     -0.0005        EQ524     -0.0005     
    X1        EQ532     -0.176         EQ538     45.5628     
    X1        EQ547     0.7302         EQ567     -0.21       
    X1        EQ577     -0.035         EQ602     -0.205      
    X1        EQ604     -0.0005        EQ614     -3E-06      
    X1        EQ636     -0.241         EQ664     -0.054      
    X1        EQ666     -0.131         EQ678     -0.003      
    X1        EQ689     -0.136         EQ690     -0.402      
    X1        EQ807     -0.0005        EQ808     -0.0005     
    X1        EQ809     -0.0005        EQ810     -0.0005     
    X1        EQ811     -0.0005        EQ871     -0.0005     
    X1        EQ872     -0.0005        EQ873     -0.0005     
    X1        EQ874     -0.0005        EQ875     -0.0005     
    X2        EQ176     -1                                   
    X3        EQ155     -1                                   
    X4        EQ156     -1                                   
    X5        EQ157     -1                                   
    X6        EQ158     -1                                   
    X7        EQ159     -1                                   
    X8        EQ160     -1                                   
    X9        EQ161     -1                                   
    X10       EQ288     -1                                   
    X11       EQ344     -1                                   
    X12       EQ459     -1                                   
    X13       EQ542     -1                                   
    X14       EQ576     -1                                   
    X15       EQ698     -1                                   
    X16       EQ699     -1                                   
    X17       EQ700     -1                                   
    X18       EQ701     -1                                   
    X19       EQ702     -1                                   
    X20       EQ703     -1                                   
    X21       EQ704     -1                                   
    X22       EQ705     -1                                   
    X23       EQ706     -1                                   
    X24       EQ707     -1                                   
    X25       EQ708     -1                                   
    X26       EQ709     -1                                   
    X27       EQ710     -1                                   
    X28       EQ711     -1                                   
    X29       EQ712     -1                                   
    X30       EQ713     -1                                   
    X31       EQ714     -1                                   
    X32       EQ715     -1                                   
    X33       EQ716     -1                                   
    X34       EQ717     -1                                   
    X35       EQ718     -1                                   
    X36       EQ719     -1                                   
    X37       EQ720     -1                                   
    X38       EQ721     -1                                   
    X39       EQ722     -1                                   
    X40       EQ723     -1                                   
    X41       EQ724     -1                                   
    X42       EQ725     -1                                   
    X43       EQ726     -1                                   
    X44       EQ727     -1                                   
    X45       EQ728     -1                                   
    X46       EQ729     -1                                   
    X47       EQ730     -1                                   
    X48       EQ731     -1                                   
    X49       EQ732     -1                                   
    X50       EQ733     -1                                   
    X51       EQ734     -1                                   
    X52       EQ735     -1                                   
    X53       EQ736     -1                                   
    X54       EQ737     -1                                   
    X55       EQ738     -1                                   
    X56       EQ739     -1                                   
    X57       EQ740     -1                                   
    X58       EQ741     -1                                   
    X59       EQ742     -1                                   
    X60       EQ743     -1                                   
    X61       EQ744     -1                                   
    X62       EQ745     -1                                   
    X63       EQ746     -1                                   
    X64       EQ747     -1                                   
    X65       EQ748     -1                                   
    X66       EQ749     -1                                   
    X67       EQ750     -1                                   
    X68       EQ751     -1                                   
    X69       EQ752     -1                                   
    X70       EQ753     -1                                   
    X71       EQ754     -1                                   
    X72       EQ755     -1                                   
    X73       EQ756     -1                                   
    X74       EQ757     -1                                   
    X75       EQ758     -1                                   
    X76       EQ759     -1                                   
    X77       EQ760     -1                                   
    X78       EQ761     -1                                   
    X79       EQ762     -1                                   
    X80       EQ763     -1                                   
    X81       EQ764     -1                                   
    X82       EQ765     -1                                   
    X83       EQ766     -1                                   
    X84       EQ767     -1                                   
    X85       EQ768     -1                                   
    X86       EQ769     -1                                   
    X87       EQ770     -1                                   
    X88       EQ771     -1                                   
    X89       EQ772     -1                                   
    X90       EQ773     -1                                   
    X91       EQ774     -1                                   
    X92       EQ775     -1                                   
    X93       EQ776     -1                                   
    X94       EQ777     -1                                   
    X95       EQ778     -1                                   
    X96       EQ779     -1                                   
    X97       EQ780     -1                                   
    X98       EQ781     -1                                   
    X99       EQ782     -1                                   
    X100      EQ783     -1                                   
    X101      EQ784     -1                                   
    X102      EQ785     -1                                   
    X103      EQ786     -1                                   
    X104      EQ145     -1                                   
    X105      EQ324     -1                                   
    X106      EQ325     -1                                   
    X107      EQ460     -1                                   
    X108      EQ552     -1                                   
    X109      EQ553     -1                                   
    X110      EQ634     -1                                   
    X111      EQ9       -1             EQ165     1           
    X111      EQ403     -1             EQ474     1           
    X111      EQ475     -1                                   
    X112      EQ25      -1             EQ150     1           
    X112      EQ201     1              EQ231     -1          
    X112      EQ403     -2                                   
    X113      EQ25      1              EQ403     1           
    X113      EQ787     -1             EQ842     -1          
    X114      EQ699     -1             EQ787     1           
    X115      EQ49      1              EQ403     1           
    X115      EQ788     -1             EQ842     -1          
    X116      EQ700     -1             EQ788     1           
    X117      EQ701     -1             EQ789     1           
    X118      EQ50      1              EQ789     -1          
    X119      EQ51      -1             EQ267     1           
    X119      EQ418     1              EQ482     -1          
    X120      EQ67      -1             EQ153     1           
    X120      EQ169     1              EQ269     -1          
    X121      EQ66      -1             EQ153     1           
    X121      EQ170     1              EQ269     -1          
    X122      EQ60      -1             EQ153     1           
    X122      EQ167     1              EQ269     -1          
    X123      EQ74      -1             EQ153     1           
    X123      EQ172     1              EQ269     -1          
    X124      EQ61      -1             EQ153     1           
    X124      EQ168     1              EQ269     -1          
    X125      EQ70      -1             EQ153     1           
    X125      EQ171     1              EQ269     -1          
    X126      EQ76      -1             EQ153     1           
    X126      EQ173     1              EQ269     -1          
    X127      EQ67      -1             EQ407     1           
    X127      EQ629     1                                    
    X128      EQ66      -1             EQ407     1           
    X128      EQ627     1                                    
    X129      EQ60      -1             EQ407     1           
    X129      EQ621     1                                    
    X130      EQ74      -1             EQ407     1           
    X130      EQ639     1                                    
    X131      EQ61      -1             EQ407     1           
    X131      EQ622     1                                    
    X132      EQ78      -1             EQ407     1           
    X132      EQ643     1                                    
    X133      EQ62      -1             EQ407     1           
    X133      EQ623     1                                    
    X134      EQ77      -1             EQ407     1           
    X134      EQ641     1                                    
    X135      EQ63      -1             EQ407     1           
    X135      EQ620     1                                    
    X136      EQ57      -1             EQ236     1           
    X136      EQ407     1                                    
    X137      EQ70      -1             EQ407     1           
    X137      EQ632     1                                    
    X138      EQ76      -1             EQ407     1           
    X138      EQ640     1                                    
    X139      EQ71      -1             EQ73      1           
    X139      EQ269     1              EQ403     1           
    X139      EQ407     -1                                   
    X140      EQ82      -1             EQ83      1           
    X140      EQ403     1              EQ472     -1          
    X140      EQ473     1                                    
    X141      EQ83      -1             EQ403     2           
    X141      EQ407     -1             EQ452     1           
    X141      EQ472     -1             EQ473     1           
    X142      EQ84      -1             EQ248     1           
    X142      EQ403     1              EQ482     -1          
    X143      EQ105     -1             EQ497     1           
    X143      EQ613     1              EQ614     -1          
    X144      EQ67      1              EQ95      -1          
    X144      EQ403     -1             EQ474     1           
    X144      EQ475     -1                                   
    X145      EQ66      1              EQ94      -1          
    X145      EQ403     -1             EQ474     1           
    X145      EQ475     -1                                   
    X146      EQ60      1              EQ88      -1          
    X146      EQ403     -1             EQ474     1           
    X146      EQ475     -1                                   
    X147      EQ74      1              EQ99      -1          
    X147      EQ403     -1             EQ474     1           
    X147      EQ475     -1                                   
    X148      EQ61      1              EQ89      -1          
    X148      EQ403     -1             EQ474     1           
    X148      EQ475     -1                                   
    X149      EQ78      1              EQ103     -1          
    X149      EQ403     -1             EQ474     1           
    X149      EQ475     -1                                   
    X150      EQ62      1              EQ90      -1          
    X150      EQ403     -1             EQ474     1           
    X150      EQ475     -1                                   
    X151      EQ77      1              EQ102     -1          
    X151      EQ403     -1             EQ474     1           
    X151      EQ475     -1                                   
    X152      EQ63      1              EQ91      -1          
    X152      EQ403     -1             EQ474     1           
    X152      EQ475     -1                                   
    X153      EQ57      1              EQ191     -1          
    X153      EQ403     -1             EQ474     1           
    X153      EQ475     -1                                   
    X154      EQ70      1              EQ98      -1          
    X154      EQ403     -1             EQ474     1           
    X154      EQ475     -1                                   
    X155      EQ76      1              EQ101     -1          
    X155      EQ403     -1             EQ474     1           
    X155      EQ475     -1                                   
    X156      EQ95      1              EQ153     1           
    X156      EQ267     1              EQ403     -1          
    X156      EQ455     -1             EQ485     -1          
    X157      EQ94      1              EQ153     1           
    X157      EQ267     1              EQ294     -1          
    X157      EQ403     -1             EQ455     -1          
    X158      EQ88      1              EQ153     1           
    X158      EQ251     -1             EQ267     1           
    X158      EQ403     -1             EQ455     -1          
    X159      EQ99      1              EQ153     1           
    X159      EQ267     1              EQ302     -1          
    X159      EQ403     -1             EQ455     -1          
    X160      EQ89      1              EQ153     1           
    X160      EQ250     -1             EQ267     1           
    X160      EQ403     -1             EQ455     -1          
    X161      EQ103     1              EQ153     1           
    X161      EQ267     1              EQ403     -1          
    X161      EQ455     -1             EQ469     -1          
    X162      EQ90      1              EQ153     1           
    X162      EQ267     1              EQ403     -1          
    X162      EQ455     -1             EQ628     -1          
    X163      EQ102     1              EQ153     1           
    X163      EQ267     1              EQ403     -1          
    X163      EQ455     -1             EQ509     -1          
    X164      EQ91      1              EQ153     1           
    X164      EQ267     1              EQ403     -1          
    X164      EQ413     -1             EQ455     -1          
    X165      EQ98      1              EQ153     1           
    X165      EQ237     -1             EQ267     1           
    X165      EQ403     -1             EQ455     -1          
    X166      EQ101     1              EQ153     1           
    X166      EQ267     1              EQ403     -1          
    X166      EQ415     -1             EQ455     -1          
    X167      EQ182     1              EQ269     -1          
    X167      EQ497     -1             EQ614     1           
    X168      EQ702     -1             EQ790     1           
    X169      EQ105     1              EQ403     1           
    X169      EQ790     -1             EQ842     -1          
    X170      EQ115     1              EQ175     -1          
    X170      EQ217     1              EQ231     -1          
    X170      EQ269     -1             EQ547     1           
    X171      EQ116     -1             EQ146     1           
    X171      EQ267     1              EQ403     -1          
    X172      EQ49      1              EQ122     -1          
    X172      EQ403     -1             EQ407     1           
    X172      EQ474     1              EQ475     -1          
    X172      EQ482     -1                                   
    X173      EQ704     -1             EQ792     1           
    X174      EQ124     1              EQ407     -1          
    X174      EQ537     -1             EQ538     1           
    X175      EQ126     -1             EQ127     1           
    X175      EQ403     1              EQ472     -1          
    X175      EQ473     1                                    
    X176      EQ127     -1             EQ403     2           
    X176      EQ407     -1             EQ472     -1          
    X176      EQ473     1              EQ501     1           
    X177      EQ128     -1             EQ403     1           
    X177      EQ425     1              EQ482     -1          
    X178      EQ267     1              EQ403     -1          
    X178      EQ493     1              EQ498     -1          
    X179      EQ29      1              EQ44      -1          
    X179      EQ267     1              EQ403     -1          
    X180      EQ30      -1             EQ498     1           
    X181      EQ44      1              EQ424     -1          
    X182      EQ149     -1             EQ267     1           
    X182      EQ318     1              EQ403     -2          
    X182      EQ407     1              EQ472     1           
    X182      EQ473     -1             EQ482     -1          
    X183      EQ164     1              EQ403     -1          
    X183      EQ407     -1             EQ635     -1          
    X184      EQ163     1              EQ164     -1          
    X184      EQ403     2              EQ407     -1          
    X184      EQ472     -1             EQ473     1           
    X185      EQ45      -1             EQ163     1           
    X185      EQ208     1              EQ378     -1          
    X186      EQ109     -1             EQ208     -1          
    X186      EQ378     1              EQ619     1           
    X187      EQ109     1              EQ403     1           
    X187      EQ791     -1             EQ842     -1          
    X188      EQ703     -1             EQ791     1           
    X189      EQ177     1              EQ182     -2          
    X189      EQ269     1                                    
    X190      EQ181     -1             EQ182     1           
    X190      EQ269     -1             EQ403     1           
    X190      EQ472     -1             EQ473     1           
    X191      EQ199     -1             EQ202     1           
    X191      EQ401     -1             EQ403     -1          
    X191      EQ547     1                                    
    X192      EQ182     -1             EQ201     1           
    X192      EQ231     -1             EQ403     1           
    X192      EQ410     -1             EQ456     1           
    X192      EQ538     1                                    
    X193      EQ183     1              EQ186     -1          
    X193      EQ201     1              EQ231     -1          
    X194      EQ182     -1             EQ186     1           
    X194      EQ269     1              EQ378     -1          
    X194      EQ403     1                                    
    X195      EQ19      1              EQ43      -1          
    X195      EQ267     1              EQ403     -1          
    X195      EQ581     -1                                   
    X196      EQ178     -1             EQ192     1           
    X196      EQ201     1              EQ231     -1          
    X197      EQ213     1              EQ267     1           
    X197      EQ403     -1             EQ581     -2          
    X198      EQ232     1              EQ235     -1          
    X198      EQ342     -1             EQ343     1           
    X199      EQ342     -1             EQ343     1           
    X199      EQ429     1              EQ431     -1          
    X200      EQ342     -1             EQ343     1           
    X200      EQ484     1              EQ487     -1          
    X201      EQ293     1              EQ295     -1          
    X201      EQ342     -1             EQ343     1           
    X202      EQ301     1              EQ303     -1          
    X202      EQ342     -1             EQ343     1           
    X203      EQ342     -1             EQ343     1           
    X203      EQ625     1              EQ626     -1          
    X204      EQ342     -1             EQ343     1           
    X204      EQ412     1              EQ539     -1          
    X205      EQ36      1              EQ37      -1          
    X205      EQ342     -1             EQ343     1           
    X206      EQ42      2              EQ407     2           
    X206      EQ433     -2             EQ482     -1          
    X207      EQ153     -1             EQ179     1           
    X207      EQ182     -1             EQ269     1           
    X208      EQ178     1              EQ189     -1          
    X208      EQ407     -1             EQ494     1           
    X209      EQ707     -1             EQ795     1           
    X210      EQ263     -1             EQ434     1           
    X211      EQ188     1              EQ263     -1          
    X211      EQ407     1                                    
    X212      EQ188     -1             EQ407     -1          
    X212      EQ434     1                                    
    X213      EQ184     1              EQ189     -1          
    X213      EQ208     -1             EQ378     1           
    X214      EQ217     1              EQ231     -1          
    X214      EQ269     -1             EQ541     -1          
    X214      EQ544     1              EQ547     1           
    X215      EQ178     -1             EQ182     1           
    X215      EQ217     1              EQ231     -1          
    X215      EQ269     -1             EQ547     1           
    X216      EQ178     1              EQ403     1           
    X216      EQ794     -1             EQ842     -1          
    X217      EQ706     -1             EQ794     1           
    X218      EQ197     -1             EQ403     -1          
    X218      EQ407     -1             EQ440     1           
    X218      EQ478     1                                    
    X219      EQ110     1              EQ111     -1          
    X219      EQ403     1              EQ581     1           
    X220      EQ194     -1             EQ198     1           
    X220      EQ201     1              EQ220     -1          
    X220      EQ231     -1             EQ403     1           
    X220      EQ538     1                                    
    X221      EQ194     -1             EQ199     1           
    X221      EQ201     1              EQ221     -1          
    X221      EQ231     -1             EQ403     1           
    X221      EQ538     1                                    
    X222      EQ193     -1             EQ194     1           
    X222      EQ201     4              EQ231     -4          
    X222      EQ375     -4             EQ378     4           
    X222      EQ403     4              EQ407     -4          
    X222      EQ538     4                                    
    X223      EQ201     2              EQ217     -1          
    X223      EQ231     -1                                   
    X224      EQ231     -1             EQ244     1           
    X224      EQ547     1                                    
    X225      EQ198     -1             EQ199     1           
    X225      EQ201     1              EQ231     -1          
    X225      EQ403     1                                    
    X226      EQ200     1              EQ202     -1          
    X226      EQ396     1              EQ403     1           
    X226      EQ591     -1                                   
    X227      EQ195     -1             EQ217     1           
    X227      EQ547     1              EQ569     -1          
    X228      EQ154     1              EQ206     -1          
    X228      EQ547     -1             EQ569     1           
    X229      EQ217     1              EQ296     -1          
    X229      EQ360     1                                    
    X230      EQ15      -1             EQ206     1           
    X230      EQ360     1                                    
    X231      EQ227     -1             EQ296     1           
    X231      EQ371     1              EQ401     -1          
    X231      EQ403     2              EQ438     -1          
    X231      EQ538     1                                    
    X232      EQ203     -1             EQ407     -1          
    X232      EQ577     1              EQ687     1           
    X233      EQ4       -1             EQ153     1           
    X233      EQ302     -1             EQ503     1           
    X234      EQ6       -1             EQ153     1           
    X234      EQ504     1              EQ509     -1          
    X235      EQ5       -1             EQ153     1           
    X235      EQ413     -1             EQ505     1           
    X236      EQ8       -1             EQ153     1           
    X236      EQ488     -1             EQ506     1           
    X237      EQ7       -1             EQ153     1           
    X237      EQ489     -1             EQ507     1           
    X238      EQ183     1              EQ184     -1          
    X238      EQ403     1              EQ474     -1          
    X238      EQ475     1              EQ538     -1          
    X239      EQ1       -1             EQ206     -1          
    X239      EQ358     1              EQ633     1           
    X240      EQ137     1              EQ207     -1          
    X240      EQ267     -1             EQ403     1           
    X241      EQ208     -1             EQ267     1           
    X241      EQ269     -1             EQ472     -1          
    X241      EQ473     1              EQ614     1           
    X242      EQ208     -1             EQ267     1           
    X242      EQ403     -1             EQ448     -1          
    X242      EQ599     1                                    
    X243      EQ269     -1             EQ313     1           
    X243      EQ599     -1             EQ614     1           
    X244      EQ208     1              EQ403     1           
    X244      EQ797     -1             EQ842     -1          
    X245      EQ708     -1             EQ797     1           
    X246      EQ201     1              EQ210     -2          
    X246      EQ212     1              EQ231     -1          
    X246      EQ403     1              EQ538     1           
    X247      EQ210     1              EQ211     -1          
    X248      EQ208     -1             EQ211     -1          
    X248      EQ378     1              EQ581     1           
    X249      EQ211     -1             EQ214     1           
    X249      EQ217     1              EQ231     -1          
    X249      EQ547     1              EQ646     -1          
    X250      EQ709     -1             EQ798     1           
    X251      EQ389     -1             EQ391     1           
    X251      EQ403     -1             EQ472     1           
    X251      EQ473     -1                                   
    X252      EQ181     -1             EQ339     1           
    X252      EQ403     -1             EQ472     1           
    X252      EQ473     -1                                   
    X253      EQ181     1              EQ339     -1          
    X253      EQ552     -1             EQ553     1           
    X254      EQ178     1              EQ181     -1          
    X254      EQ403     2              EQ407     -1          
    X254      EQ472     -1             EQ473     1           
    X255      EQ178     1              EQ181     -1          
    X255      EQ403     2              EQ407     -1          
    X255      EQ472     -1             EQ473     1           
    X256      EQ113     1              EQ207     -1          
    X256      EQ355     2              EQ403     3           
    X256      EQ407     -1             EQ472     -1          
    X256      EQ473     1                                    
    X257      EQ195     1              EQ217     -1          
    X257      EQ407     -1             EQ588     1           
    X258      EQ218     -1             EQ547     1           
    X258      EQ555     1              EQ569     -1          
    X259      EQ218     1              EQ261     -1          
    X259      EQ375     -1             EQ378     1           
    X259      EQ403     1              EQ581     1           
    X260      EQ9       1              EQ148     -1          
    X260      EQ407     1                                    
    X261      EQ201     2              EQ219     -1          
    X261      EQ403     2              EQ407     -1          
    X262      EQ139     -1             EQ142     1           
    X262      EQ407     -1             EQ477     1           
    X263      EQ140     1              EQ141     -1          
    X263      EQ403     -1             EQ474     1           
    X263      EQ475     -1                                   
    X264      EQ217     1              EQ222     -1          
    X264      EQ403     2              EQ609     1           
    X264      EQ644     1              EQ645     -1          
    X265      EQ142     -1             EQ208     -1          
    X265      EQ378     1              EQ500     1           
    X266      EQ223     -1             EQ264     1           
    X266      EQ407     -1             EQ478     1           
    X267      EQ203     1              EQ223     -1          
    X267      EQ267     1              EQ403     -1          
    X268      EQ223     1              EQ494     -1          
    X268      EQ800     -1             EQ868     1           
    X269      EQ223     1              EQ224     -1          
    X269      EQ360     1                                    
    X270      EQ217     1              EQ224     1           
    X270      EQ227     -1             EQ231     -1          
    X270      EQ264     -1             EQ403     1           
    X270      EQ547     1                                    
    X271      EQ217     1              EQ223     -1          
    X271      EQ225     1              EQ231     -1          
    X271      EQ547     1              EQ647     -1          
    X272      EQ201     1              EQ223     1           
    X272      EQ231     -1             EQ403     1           
    X272      EQ407     -1             EQ538     1           
    X272      EQ800     -1                                   
    X273      EQ710     -1             EQ800     1           
    X274      EQ211     1              EQ223     -1          
    X274      EQ499     1              EQ581     -1          
    X275      EQ131     -1             EQ228     1           
    X275      EQ403     -1             EQ474     1           
    X275      EQ475     -1             EQ538     1           
    X276      EQ226     -1             EQ227     1           
    X276      EQ407     -1             EQ478     1           
    X277      EQ217     1              EQ226     1           
    X277      EQ227     -1             EQ231     -1          
    X277      EQ375     -1             EQ378     1           
    X277      EQ403     1              EQ407     -1          
    X277      EQ547     1                                    
    X278      EQ227     -1             EQ246     1           
    X278      EQ247     -1             EQ403     1           
    X278      EQ538     1                                    
    X279      EQ131     1              EQ201     1           
    X279      EQ227     -1             EQ231     -1          
    X280      EQ227     -1             EQ407     -1          
    X280      EQ408     1              EQ478     1           
    X280      EQ482     -1             EQ483     1           
    X281      EQ227     -1             EQ403     1           
    X281      EQ408     1              EQ432     1           
    X281      EQ482     -1                                   
    X282      EQ227     -1             EQ360     1           
    X282      EQ478     1                                    
    X283      EQ208     -1             EQ227     -1          
    X283      EQ378     1              EQ483     1           
    X284      EQ217     1              EQ227     -1          
    X284      EQ229     1              EQ231     -1          
    X284      EQ547     1              EQ648     -1          
    X285      EQ201     1              EQ227     1           
    X285      EQ231     -1             EQ403     1           
    X285      EQ407     -1             EQ538     1           
    X285      EQ801     -1                                   
    X286      EQ227     1              EQ403     2           
    X286      EQ801     -1             EQ842     -2          
    X287      EQ711     -1             EQ801     1           
    X288      EQ223     -1             EQ269     1           
    X288      EQ403     1              EQ612     1           
    X288      EQ614     -1                                   
    X289      EQ231     -1             EQ547     1           
    X289      EQ557     1              EQ569     -1          
    X290      EQ201     -1             EQ231     1           
    X290      EQ403     3              EQ407     1           
    X290      EQ538     -1             EQ842     -4          
    X291      EQ180     1              EQ234     -1          
    X291      EQ403     1              EQ472     -1          
    X291      EQ473     1                                    
    X292      EQ233     -1             EQ390     1           
    X292      EQ403     2              EQ407     -1          
    X292      EQ472     -1             EQ473     1           
    X293      EQ321     1              EQ390     -1          
    X293      EQ411     -1             EQ461     1           
    X294      EQ238     -1             EQ239     1           
    X294      EQ403     -1             EQ472     1           
    X294      EQ473     -1             EQ482     -1          
    X295      EQ240     1              EQ241     -1          
    X295      EQ403     1              EQ472     -1          
    X295      EQ473     1                                    
    X296      EQ238     1              EQ240     -1          
    X296      EQ403     2              EQ407     -1          
    X296      EQ472     -1             EQ473     1           
    X297      EQ239     -1             EQ245     1           
    X297      EQ267     1              EQ472     -1          
    X297      EQ473     1                                    
    X298      EQ238     1              EQ403     1           
    X298      EQ802     -1             EQ842     -1          
    X299      EQ712     -1             EQ802     1           
    X300      EQ48      1              EQ182     1           
    X300      EQ242     -1             EQ407     -1          
    X301      EQ50      -1             EQ217     1           
    X301      EQ231     -1             EQ242     1           
    X301      EQ269     -1             EQ547     1           
    X302      EQ407     2              EQ408     -2          
    X302      EQ482     1                                    
    X303      EQ31      1              EQ245     -1          
    X303      EQ482     -1                                   
    X304      EQ245     -1             EQ249     1           
    X304      EQ403     2              EQ482     -1          
    X305      EQ713     -1             EQ803     1           
    X306      EQ245     1              EQ803     -1          
    X307      EQ201     1              EQ231     -1          
    X307      EQ247     1              EQ267     -1          
    X307      EQ403     2              EQ478     -1          
    X308      EQ201     2              EQ231     -2          
    X308      EQ247     1              EQ375     -1          
    X308      EQ378     1              EQ403     2           
    X308      EQ407     -1             EQ410     -1          
    X308      EQ538     1                                    
    X309      EQ46      1              EQ114     -1          
    X309      EQ201     1              EQ231     -1          
    X309      EQ403     1                                    
    X310      EQ204     2              EQ215     -2          
    X310      EQ273     1              EQ403     2           
    X310      EQ515     -1                                   
    X311      EQ204     2              EQ215     -2          
    X311      EQ275     1              EQ403     2           
    X311      EQ523     -1                                   
    X312      EQ204     2              EQ215     -2          
    X312      EQ274     1              EQ403     2           
    X312      EQ517     -1                                   
    X313      EQ204     2              EQ215     -2          
    X313      EQ276     1              EQ403     2           
    X313      EQ525     -1                                   
    X314      EQ201     1              EQ231     -1          
    X314      EQ259     1              EQ403     1           
    X314      EQ407     -1             EQ538     1           
    X314      EQ804     -1                                   
    X315      EQ259     1              EQ403     1           
    X315      EQ804     -1             EQ842     -1          
    X316      EQ714     -1             EQ804     1           
    X317      EQ233     1              EQ259     -1          
    X317      EQ403     1              EQ472     -1          
    X317      EQ473     1                                    
    X318      EQ715     -1             EQ805     1           
    X319      EQ259     1              EQ260     -1          
    X319      EQ403     1              EQ407     -1          
    X319      EQ610     1                                    
    X320      EQ201     1              EQ231     -1          
    X320      EQ260     1              EQ403     1           
    X320      EQ407     -1             EQ538     1           
    X320      EQ805     -1                                   
    X321      EQ261     -1             EQ546     1           
    X322      EQ108     -1             EQ261     1           
    X322      EQ538     1                                    
    X323      EQ263     1              EQ403     1           
    X323      EQ806     -1             EQ842     -1          
    X324      EQ716     -1             EQ806     1           
    X325      EQ807     1              EQ838     1           
    X325      EQ877     -2                                   
    X326      EQ808     1              EQ838     1           
    X326      EQ878     -2                                   
    X327      EQ809     1              EQ838     1           
    X327      EQ879     -2                                   
    X328      EQ810     1              EQ838     1           
    X328      EQ880     -2                                   
    X329      EQ811     1              EQ838     1           
    X329      EQ881     -2                                   
    X330      EQ717     -1             EQ813     1           
    X331      EQ267     1              EQ813     -1          
    X332      EQ271     -1             EQ272     1           
    X332      EQ403     1              EQ472     -1          
    X332      EQ473     1                                    
    X333      EQ272     -1             EQ347     1           
    X333      EQ403     2              EQ407     -1          
    X333      EQ472     -1             EQ473     1           
    X334      EQ718     -1             EQ814     1           
    X335      EQ270     1              EQ814     -1          
    X336      EQ201     1              EQ231     -1          
    X336      EQ268     1              EQ270     -1          
    X336      EQ403     2              EQ407     -1          
    X336      EQ416     -1             EQ538     1           
    X337      EQ115     -1             EQ121     1           
    X337      EQ182     1              EQ407     -1          
    X338      EQ719     -1             EQ815     1           
    X339      EQ271     1              EQ815     -1          
    X340      EQ267     2              EQ278     -1          
    X340      EQ403     -2             EQ407     2           
    X340      EQ482     -1             EQ549     1           
    X341      EQ215     -2             EQ267     2           
    X341      EQ278     -1             EQ288     2           
    X341      EQ461     2              EQ549     1           
    X342      EQ201     1              EQ231     -1          
    X342      EQ279     1              EQ403     1           
    X342      EQ407     -1             EQ538     1           
    X342      EQ816     -1                                   
    X343      EQ720     -1             EQ816     1           
    X344      EQ182     -1             EQ263     1           
    X344      EQ269     1              EQ403     1           
    X344      EQ407     -1             EQ483     -1          
    X345      EQ280     -1             EQ403     -1          
    X345      EQ407     -1             EQ478     1           
    X345      EQ684     1                                    
    X346      EQ201     1              EQ231     -1          
    X346      EQ281     1              EQ375     -1          
    X346      EQ378     1              EQ403     2           
    X346      EQ407     -1             EQ538     1           
    X346      EQ689     -1                                   
    X347      EQ350     -2             EQ354     2           
    X347      EQ403     -2             EQ582     1           
    X347      EQ583     -1             EQ842     2           
    X348      EQ350     2              EQ354     -2          
    X348      EQ407     1              EQ482     -0.5        
    X349      EQ193     1              EQ231     -1          
    X349      EQ266     -1             EQ403     -1          
    X349      EQ550     1                                    
    X350      EQ266     2              EQ268     -2          
    X350      EQ403     1              EQ472     1           
    X350      EQ473     -1                                   
    X351      EQ178     1              EQ190     -1          
    X351      EQ282     1              EQ403     1           
    X351      EQ409     -1                                   
    X352      EQ178     1              EQ190     -1          
    X352      EQ598     1              EQ665     -1          
    X353      EQ43      1              EQ282     1           
    X353      EQ283     -1             EQ407     -1          
    X353      EQ478     1                                    
    X354      EQ283     -1             EQ407     -1          
    X354      EQ411     1              EQ478     1           
    X354      EQ581     1                                    
    X355      EQ217     1              EQ231     -1          
    X355      EQ282     -1             EQ284     1           
    X355      EQ547     1              EQ649     -1          
    X356      EQ201     1              EQ231     -1          
    X356      EQ282     1              EQ403     1           
    X356      EQ407     -1             EQ538     1           
    X356      EQ817     -1                                   
    X357      EQ721     -1             EQ817     1           
    X358      EQ403     -2             EQ407     1           
    X358      EQ482     -0.5           EQ582     1           
    X358      EQ583     -1             EQ842     2           
    X359      EQ403     -4             EQ407     1           
    X359      EQ482     -0.5           EQ582     1           
    X359      EQ583     -1             EQ842     4           
    X360      EQ201     1              EQ231     -1          
    X360      EQ252     1              EQ265     -1          
    X361      EQ201     1              EQ231     -1          
    X361      EQ297     1              EQ298     -1          
    X362      EQ403     1              EQ444     1           
    X362      EQ842     -1             EQ853     -1          
    X363      EQ748     -1             EQ853     1           
    X364      EQ210     -1             EQ342     -1          
    X364      EQ343     1              EQ407     -1          
    X364      EQ478     1              EQ581     1           
    X365      EQ14      -1             EQ208     -1          
    X365      EQ228     1              EQ378     1           
    X365      EQ403     -2                                   
    X366      EQ287     -1             EQ319     1           
    X366      EQ403     -1             EQ407     -1          
    X366      EQ478     1                                    
    X367      EQ201     1              EQ231     -1          
    X367      EQ289     1              EQ290     -1          
    X368      EQ3       -1             EQ208     -1          
    X368      EQ378     1              EQ403     1           
    X368      EQ407     1              EQ551     1           
    X369      EQ18      -1             EQ267     1           
    X369      EQ403     -1             EQ450     1           
    X370      EQ17      -1             EQ18      1           
    X371      EQ253     1              EQ281     -1          
    X371      EQ403     -1             EQ503     -1          
    X371      EQ547     1                                    
    X372      EQ255     1              EQ281     -1          
    X372      EQ403     -1             EQ504     -1          
    X372      EQ547     1                                    
    X373      EQ254     1              EQ281     -1          
    X373      EQ403     -1             EQ505     -1          
    X373      EQ547     1                                    
    X374      EQ257     1              EQ281     -1          
    X374      EQ403     -1             EQ506     -1          
    X374      EQ547     1                                    
    X375      EQ256     1              EQ281     -1          
    X375      EQ403     -1             EQ507     -1          
    X375      EQ547     1                                    
    X376      EQ292     1              EQ355     1           
    X376      EQ403     1              EQ594     -1          
    X377      EQ722     -1             EQ821     1           
    X378      EQ299     -1             EQ334     1           
    X378      EQ403     -1             EQ407     -1          
    X378      EQ478     1                                    
    X379      EQ281     -1             EQ403     -1          
    X379      EQ407     -1             EQ478     1           
    X379      EQ689     1                                    
    X380      EQ723     -1             EQ824     1           
    X381      EQ23      1              EQ336     -1          
    X381      EQ407     -1             EQ519     -1          
    X381      EQ538     1                                    
    X382      EQ201     1              EQ231     -1          
    X382      EQ304     1              EQ305     -1          
    X383      EQ84      1              EQ267     1           
    X383      EQ308     -1             EQ472     -1          
    X383      EQ473     1                                    
    X384      EQ128     1              EQ267     1           
    X384      EQ309     -1             EQ472     -1          
    X384      EQ473     1                                    
    X385      EQ12      -1             EQ86      1           
    X385      EQ407     1                                    
    X386      EQ13      -1             EQ87      1           
    X386      EQ407     1                                    
    X387      EQ11      -1             EQ403     -1          
    X387      EQ474     1              EQ475     -1          
    X387      EQ631     1                                    
    X388      EQ11      1              EQ228     -1          
    X388      EQ403     1              EQ407     2           
    X388      EQ581     -1                                   
    X389      EQ312     -1             EQ403     -1          
    X389      EQ474     1              EQ475     -1          
    X389      EQ633     1                                    
    X390      EQ201     1              EQ231     -1          
    X390      EQ312     1              EQ317     -1          
    X390      EQ378     -1             EQ403     1           
    X390      EQ538     1                                    
    X391      EQ21      1              EQ314     -1          
    X391      EQ369     1              EQ403     1           
    X392      EQ315     -1             EQ408     1           
    X392      EQ482     -1             EQ495     1           
    X393      EQ246     1              EQ315     -1          
    X393      EQ403     1              EQ407     -1          
    X394      EQ27      -1             EQ58      1           
    X394      EQ182     1              EQ269     -1          
    X394      EQ403     1              EQ404     1           
    X394      EQ407     -2                                   
    X395      EQ136     -1             EQ243     1           
    X395      EQ403     1              EQ407     -1          
    X396      EQ16      -1             EQ141     1           
    X396      EQ403     -1             EQ407     -1          
    X396      EQ478     1                                    
    X397      EQ311     -1             EQ403     -1          
    X397      EQ472     1              EQ473     -1          
    X397      EQ630     1                                    
    X398      EQ21      -1             EQ110     -1          
    X398      EQ317     1              EQ407     1           
    X399      EQ55      -1             EQ56      1           
    X399      EQ407     1                                    
    X400      EQ23      -1             EQ55      1           
    X400      EQ538     1                                    
    X401      EQ49      1              EQ56      -1          
    X401      EQ407     1                                    
    X402      EQ148     1              EQ208     -1          
    X402      EQ377     1              EQ449     -1          
    X403      EQ148     1              EQ210     1           
    X403      EQ449     -1             EQ581     -1          
    X404      EQ323     -1             EQ397     1           
    X404      EQ441     -1             EQ547     1           
    X405      EQ321     -1             EQ342     -1          
    X405      EQ343     1              EQ344     1           
    X405      EQ407     -1             EQ596     1           
    X406      EQ204     1              EQ215     -1          
    X406      EQ324     -1             EQ325     1           
    X406      EQ403     1                                    
    X407      EQ314     1              EQ316     -1          
    X407      EQ407     -1             EQ538     1           
    X408      EQ205     -1             EQ316     1           
    X408      EQ403     1              EQ407     -1          
    X408      EQ547     1                                    
    X409      EQ201     1              EQ231     -1          
    X409      EQ269     1              EQ327     -1          
    X409      EQ403     1                                    
    X410      EQ155     -1             EQ169     1           
    X410      EQ269     -1             EQ460     1           
    X411      EQ156     -1             EQ170     1           
    X411      EQ269     -1             EQ460     1           
    X412      EQ157     -1             EQ167     1           
    X412      EQ269     -1             EQ460     1           
    X413      EQ158     -1             EQ172     1           
    X413      EQ269     -1             EQ460     1           
    X414      EQ159     -1             EQ168     1           
    X414      EQ269     -1             EQ460     1           
    X415      EQ160     -1             EQ171     1           
    X415      EQ269     -1             EQ460     1           
    X416      EQ161     -1             EQ173     1           
    X416      EQ269     -1             EQ460     1           
    X417      EQ26      -1             EQ403     -1          
    X417      EQ474     1              EQ475     -1          
    X417      EQ511     1                                    
    X418      EQ201     1              EQ231     -1          
    X418      EQ329     1              EQ330     -1          
    X419      EQ136     -1             EQ403     1           
    X419      EQ474     -1             EQ475     1           
    X419      EQ684     1                                    
    X420      EQ333     1              EQ334     -1          
    X420      EQ403     1              EQ407     -1          
    X420      EQ547     1                                    
    X421      EQ40      1              EQ335     -1          
    X421      EQ403     -1             EQ474     1           
    X421      EQ475     -1                                   
    X422      EQ267     1              EQ335     1           
    X422      EQ362     -1             EQ403     -1          
    X422      EQ581     -1                                   
    X423      EQ132     1              EQ336     -1          
    X423      EQ403     2              EQ407     -1          
    X423      EQ472     -1             EQ473     1           
    X424      EQ294     1              EQ403     -1          
    X424      EQ474     1              EQ475     -1          
    X424      EQ629     -1                                   
    X425      EQ302     1              EQ403     -1          
    X425      EQ474     1              EQ475     -1          
    X425      EQ627     -1                                   
    X426      EQ250     1              EQ403     -1          
    X426      EQ474     1              EQ475     -1          
    X426      EQ621     -1                                   
    X427      EQ403     -1             EQ469     1           
    X427      EQ474     1              EQ475     -1          
    X427      EQ639     -1                                   
    X428      EQ403     -1             EQ474     1           
    X428      EQ475     -1             EQ622     -1          
    X428      EQ628     1                                    
    X429      EQ403     -1             EQ474     1           
    X429      EQ475     -1             EQ509     1           
    X429      EQ643     -1                                   
    X430      EQ403     -1             EQ413     1           
    X430      EQ474     1              EQ475     -1          
    X430      EQ623     -1                                   
    X431      EQ403     -1             EQ474     1           
    X431      EQ475     -1             EQ488     1           
    X431      EQ641     -1                                   
    X432      EQ403     -1             EQ474     1           
    X432      EQ475     -1             EQ489     1           
    X432      EQ620     -1                                   
    X433      EQ236     -1             EQ237     1           
    X433      EQ403     -1             EQ474     1           
    X433      EQ475     -1                                   
    X434      EQ403     -1             EQ415     1           
    X434      EQ474     1              EQ475     -1          
    X434      EQ632     -1                                   
    X435      EQ403     -1             EQ474     1           
    X435      EQ475     -1             EQ485     1           
    X435      EQ640     -1                                   
    X436      EQ59      -1             EQ232     1           
    X436      EQ407     1                                    
    X437      EQ42      -1             EQ71      1           
    X437      EQ407     -1                                   
    X438      EQ68      -1             EQ407     1           
    X438      EQ429     1                                    
    X439      EQ75      -1             EQ407     1           
    X439      EQ484     1                                    
    X440      EQ64      -1             EQ293     1           
    X440      EQ407     1                                    
    X441      EQ65      -1             EQ301     1           
    X441      EQ407     1                                    
    X442      EQ79      -1             EQ407     1           
    X442      EQ625     1                                    
    X443      EQ69      -1             EQ407     1           
    X443      EQ412     1                                    
    X444      EQ36      -1             EQ72      1           
    X444      EQ407     -1                                   
    X445      EQ24      -1             EQ362     1           
    X445      EQ581     1                                    
    X446      EQ24      1              EQ151     -1          
    X446      EQ407     1                                    
    X447      EQ47      -1             EQ407     1           
    X447      EQ519     1                                    
    X448      EQ201     1              EQ231     -1          
    X448      EQ338     1              EQ403     1           
    X448      EQ407     -1             EQ538     1           
    X448      EQ828     -1                                   
    X449      EQ724     -1             EQ828     1           
    X450      EQ217     1              EQ231     -1          
    X450      EQ269     -1             EQ295     1           
    X450      EQ403     1              EQ547     1           
    X450      EQ821     -1             EQ842     -1          
    X451      EQ217     1              EQ231     -1          
    X451      EQ269     -1             EQ303     1           
    X451      EQ403     1              EQ547     1           
    X451      EQ824     -1             EQ842     -1          
    X452      EQ217     1              EQ231     -1          
    X452      EQ269     -1             EQ403     1           
    X452      EQ547     1              EQ626     1           
    X452      EQ842     -1             EQ903     -1          
    X453      EQ217     1              EQ231     -1          
    X453      EQ269     -1             EQ403     1           
    X453      EQ539     1              EQ547     1           
    X453      EQ842     -1             EQ845     -1          
    X454      EQ217     1              EQ231     -1          
    X454      EQ269     -1             EQ403     1           
    X454      EQ431     1              EQ547     1           
    X454      EQ842     -1             EQ847     -1          
    X455      EQ217     1              EQ231     -1          
    X455      EQ269     -1             EQ403     1           
    X455      EQ487     1              EQ547     1           
    X455      EQ842     -1             EQ867     -1          
    X456      EQ285     2              EQ344     -2          
    X456      EQ354     -2             EQ355     2           
    X456      EQ403     4              EQ482     -1          
    X457      EQ310     1              EQ345     -1          
    X457      EQ362     1                                    
    X458      EQ341     1              EQ345     -1          
    X458      EQ407     -1             EQ538     1           
    X459      EQ346     -1             EQ403     2           
    X459      EQ533     1              EQ548     -1          
    X460      EQ182     1              EQ348     -1          
    X460      EQ407     -1             EQ692     1           
    X461      EQ267     1              EQ355     -1          
    X461      EQ472     -1             EQ473     1           
    X462      EQ352     1              EQ353     -1          
    X462      EQ369     1              EQ403     1           
    X462      EQ407     1              EQ442     -1          
    X462      EQ482     -1             EQ609     1           
    X463      EQ344     1              EQ352     1           
    X463      EQ353     -1             EQ403     1           
    X463      EQ407     1              EQ467     -1          
    X463      EQ482     -1             EQ609     1           
    X464      EQ181     1              EQ338     -1          
    X464      EQ352     1              EQ353     -1          
    X464      EQ403     1              EQ407     1           
    X464      EQ482     -1             EQ609     1           
    X465      EQ352     1              EQ353     -1          
    X465      EQ403     1              EQ407     1           
    X465      EQ482     -1             EQ518     -1          
    X465      EQ576     1              EQ609     1           
    X466      EQ176     1              EQ352     1           
    X466      EQ353     -1             EQ403     1           
    X466      EQ407     1              EQ482     -1          
    X466      EQ609     1              EQ624     -1          
    X467      EQ201     1              EQ231     -1          
    X467      EQ346     1              EQ403     1           
    X467      EQ407     -1             EQ538     1           
    X467      EQ725     -1                                   
    X468      EQ346     1              EQ725     -1          
    X469      EQ217     1              EQ231     -1          
    X469      EQ269     -1             EQ347     -1          
    X469      EQ348     1              EQ547     1           
    X470      EQ726     -1             EQ829     1           
    X471      EQ347     1              EQ829     -1          
    X472      EQ1       -1             EQ351     1           
    X472      EQ403     1              EQ463     -1          
    X472      EQ633     1                                    
    X473      EQ231     -1             EQ342     1           
    X473      EQ352     -1             EQ403     -1          
    X473      EQ547     1                                    
    X474      EQ352     -1             EQ353     1           
    X474      EQ403     -1             EQ474     1           
    X474      EQ475     -1                                   
    X475      EQ20      -1             EQ110     -1          
    X475      EQ317     1              EQ547     1           
    X476      EQ166     1              EQ356     -1          
    X476      EQ403     -2             EQ407     -1          
    X476      EQ477     1                                    
    X477      EQ201     1              EQ231     -1          
    X477      EQ340     -1             EQ345     1           
    X477      EQ403     1                                    
    X478      EQ340     1              EQ519     -1          
    X478      EQ581     1              EQ727     -1          
    X479      EQ1       -1             EQ355     1           
    X479      EQ403     1              EQ407     -1          
    X479      EQ633     1                                    
    X480      EQ360     -1             EQ407     -1          
    X480      EQ454     1                                    
    X481      EQ117     -1             EQ180     1           
    X481      EQ360     1              EQ403     1           
    X481      EQ407     -1                                   
    X482      EQ360     1              EQ403     2           
    X482      EQ830     -1             EQ842     -2          
    X483      EQ728     -1             EQ830     1           
    X484      EQ182     -1             EQ185     1           
    X484      EQ269     1              EQ366     -1          
    X484      EQ403     1                                    
    X485      EQ138     1              EQ379     -1          
    X486      EQ4       1              EQ153     1           
    X486      EQ302     -1             EQ393     -1          
    X487      EQ6       1              EQ153     1           
    X487      EQ393     -1             EQ509     -1          
    X488      EQ5       1              EQ153     1           
    X488      EQ393     -1             EQ413     -1          
    X489      EQ8       1              EQ153     1           
    X489      EQ393     -1             EQ488     -1          
    X490      EQ7       1              EQ153     1           
    X490      EQ393     -1             EQ489     -1          
    X491      EQ310     1              EQ342     -1          
    X491      EQ343     1              EQ393     -1          
    X492      EQ310     1              EQ393     -1          
    X492      EQ403     1              EQ474     -1          
    X492      EQ475     1                                    
    X493      EQ10      1              EQ381     -1          
    X493      EQ403     1              EQ407     1           
    X494      EQ380     -1             EQ381     1           
    X494      EQ403     -1             EQ474     1           
    X494      EQ475     -1             EQ538     1           
    X495      EQ152     1              EQ364     -1          
    X495      EQ403     1              EQ474     -1          
    X495      EQ475     1                                    
    X496      EQ152     1              EQ363     -1          
    X496      EQ403     1              EQ474     -1          
    X496      EQ475     1                                    
    X497      EQ363     -1             EQ364     1           
    X498      EQ365     -1             EQ403     2           
    X498      EQ482     -1             EQ492     1           
    X499      EQ342     -1             EQ343     1           
    X499      EQ787     1              EQ833     -1          
    X500      EQ361     -1             EQ403     -1          
    X500      EQ547     1              EQ678     1           
    X500      EQ689     -1                                   
    X501      EQ2       1              EQ362     -1          
    X501      EQ403     1              EQ472     -1          
    X501      EQ473     1              EQ538     -1          
    X502      EQ1       -1             EQ349     1           
    X502      EQ368     -1             EQ403     1           
    X502      EQ633     1                                    
    X503      EQ201     1              EQ231     -1          
    X503      EQ349     1              EQ355     -1          
    X503      EQ368     -1             EQ403     1           
    X503      EQ538     1                                    
    X504      EQ729     -1             EQ831     1           
    X505      EQ365     1              EQ403     1           
    X505      EQ831     -1             EQ842     -1          
    X506      EQ369     -1             EQ394     1           
    X506      EQ403     2              EQ407     -1          
    X506      EQ472     -1             EQ473     1           
    X507      EQ118     1              EQ370     -1          
    X507      EQ403     2              EQ407     -1          
    X507      EQ472     -1             EQ473     1           
    X508      EQ372     -1             EQ403     1           
    X508      EQ407     -1             EQ538     1           
    X508      EQ545     1                                    
    X509      EQ341     -1             EQ367     1           
    X509      EQ375     -1             EQ378     1           
    X510      EQ388     1              EQ407     1           
    X510      EQ465     1              EQ602     -1          
    X510      EQ633     -1                                   
    X511      EQ201     1              EQ231     -1          
    X511      EQ371     1              EQ396     -1          
    X512      EQ582     -1             EQ583     1           
    X512      EQ832     -1             EQ833     1           
    X512      EQ842     1              EQ844     -1          
    X513      EQ374     1              EQ403     1           
    X513      EQ833     -1             EQ842     -1          
    X514      EQ731     -1             EQ833     1           
    X515      EQ201     1              EQ231     -1          
    X515      EQ269     -1             EQ383     -1          
    X515      EQ384     1              EQ538     1           
    X516      EQ174     -1             EQ269     -1          
    X516      EQ313     1              EQ384     1           
    X517      EQ201     1              EQ231     -1          
    X517      EQ373     1              EQ403     1           
    X517      EQ407     -1             EQ538     1           
    X517      EQ832     -1                                   
    X518      EQ730     -1             EQ832     1           
    X519      EQ201     1              EQ231     -1          
    X519      EQ375     1              EQ378     -1          
    X519      EQ403     1              EQ478     -1          
    X519      EQ538     1                                    
    X520      EQ201     1              EQ231     -1          
    X520      EQ375     -1             EQ376     1           
    X520      EQ378     1              EQ386     -1          
    X520      EQ403     1              EQ407     -1          
    X520      EQ538     1                                    
    X521      EQ217     1              EQ231     -1          
    X521      EQ375     -1             EQ376     1           
    X521      EQ547     1              EQ650     -1          
    X522      EQ201     1              EQ231     -1          
    X522      EQ378     -1             EQ380     1           
    X523      EQ201     1              EQ231     -1          
    X523      EQ282     -1             EQ378     -1          
    X523      EQ382     1              EQ403     1           
    X523      EQ538     1                                    
    X524      EQ109     1              EQ267     1           
    X524      EQ378     -1             EQ403     -1          
    X525      EQ208     1              EQ378     -1          
    X525      EQ403     1              EQ407     -1          
    X525      EQ474     -1             EQ475     1           
    X525      EQ478     1                                    
    X526      EQ375     -1             EQ378     1           
    X526      EQ407     -1             EQ478     1           
    X527      EQ375     -1             EQ378     1           
    X527      EQ407     -1             EQ547     1           
    X527      EQ554     1              EQ569     -1          
    X528      EQ377     -1             EQ378     1           
    X529      EQ208     -1             EQ375     -1          
    X529      EQ378     2              EQ403     -1          
    X529      EQ474     1              EQ475     -1          
    X530      EQ232     1              EQ267     1           
    X530      EQ342     -1             EQ343     1           
    X530      EQ384     -1             EQ403     -1          
    X531      EQ379     1              EQ385     -1          
    X531      EQ403     -1             EQ474     1           
    X531      EQ475     -1             EQ651     1           
    X532      EQ217     1              EQ231     -1          
    X532      EQ378     -1             EQ385     1           
    X532      EQ547     1              EQ651     -1          
    X533      EQ201     1              EQ231     -1          
    X533      EQ378     1              EQ403     1           
    X533      EQ407     -1             EQ538     1           
    X533      EQ834     -1                                   
    X534      EQ378     1              EQ470     1           
    X534      EQ834     -1             EQ861     -1          
    X535      EQ732     -1             EQ834     1           
    X536      EQ28      1              EQ267     1           
    X536      EQ387     -2             EQ403     -1          
    X537      EQ389     -1             EQ392     1           
    X537      EQ403     2              EQ407     -1          
    X537      EQ472     -1             EQ473     1           
    X538      EQ734     -1             EQ836     1           
    X539      EQ389     1              EQ836     -1          
    X540      EQ201     1              EQ231     -1          
    X540      EQ390     1              EQ403     1           
    X540      EQ407     -1             EQ538     1           
    X540      EQ837     -1                                   
    X541      EQ735     -1             EQ837     1           
    X542      EQ106     1              EQ201     1           
    X542      EQ231     -1             EQ392     -1          
    X542      EQ403     1                                    
    X543      EQ267     1              EQ388     -1          
    X543      EQ465     1              EQ472     -1          
    X543      EQ473     1              EQ478     1           
    X543      EQ633     -1                                   
    X544      EQ394     1              EQ403     1           
    X544      EQ839     -1             EQ842     -1          
    X545      EQ737     -1             EQ839     1           
    X546      EQ267     -1             EQ388     1           
    X546      EQ465     -1             EQ472     1           
    X546      EQ473     -1             EQ478     -1          
    X546      EQ633     1                                    
    X547      EQ387     1              EQ394     -1          
    X547      EQ408     1              EQ482     -1          
    X548      EQ736     -1             EQ838     1           
    X549      EQ391     -1             EQ838     1           
    X550      EQ392     1              EQ403     -1          
    X550      EQ428     -1             EQ472     1           
    X550      EQ473     -1                                   
    X551      EQ201     1              EQ231     -1          
    X551      EQ391     -1             EQ393     1           
    X551      EQ403     1                                    
    X552      EQ217     1              EQ231     -1          
    X552      EQ388     -1             EQ395     1           
    X552      EQ547     1              EQ652     -1          
    X553      EQ201     1              EQ231     -1          
    X553      EQ388     1              EQ403     1           
    X553      EQ407     -1             EQ538     1           
    X553      EQ835     -1                                   
    X554      EQ733     -1             EQ835     1           
    X555      EQ217     1              EQ231     -1          
    X555      EQ375     -1             EQ378     1           
    X555      EQ396     1              EQ403     2           
    X555      EQ407     -1             EQ547     1           
    X555      EQ695     -1                                   
    X556      EQ151     -1             EQ267     1           
    X556      EQ474     -1             EQ475     1           
    X556      EQ594     1                                    
    X557      EQ151     1              EQ201     1           
    X557      EQ231     -1             EQ374     -1          
    X557      EQ403     1                                    
    X558      EQ398     -1             EQ402     1           
    X558      EQ407     -1             EQ592     1           
    X559      EQ359     1              EQ397     -1          
    X559      EQ441     -1             EQ547     1           
    X560      EQ399     -1             EQ400     2           
    X560      EQ403     -1             EQ474     1           
    X560      EQ475     -1                                   
    X561      EQ399     1              EQ400     -2          
    X561      EQ407     2              EQ408     -1          
    X562      EQ201     1              EQ231     -1          
    X562      EQ382     -1             EQ388     -1          
    X562      EQ400     1              EQ403     1           
    X562      EQ538     1                                    
    X563      EQ205     1              EQ355     1           
    X563      EQ401     -1             EQ403     1           
    X563      EQ407     -1                                   
    X564      EQ16      1              EQ355     1           
    X564      EQ401     -1             EQ403     2           
    X564      EQ407     -3             EQ547     1           
    X565      EQ217     1              EQ231     -1          
    X565      EQ372     1              EQ401     -1          
    X565      EQ403     1                                    
    X566      EQ402     -1             EQ403     -1          
    X566      EQ407     -1             EQ478     1           
    X566      EQ694     1                                    
    X567      EQ396     1              EQ402     -1          
    X567      EQ547     1              EQ569     -1          
    X568      EQ109     1              EQ118     -1          
    X568      EQ407     -1             EQ687     1           
    X569      EQ267     -1             EQ405     1           
    X569      EQ407     -1                                   
    X570      EQ403     -1             EQ405     1           
    X570      EQ410     -1                                   
    X571      EQ740     -1             EQ844     1           
    X572      EQ407     1              EQ844     -1          
    X573      EQ739     -1             EQ843     1           
    X574      EQ404     1              EQ843     -1          
    X575      EQ59      -1             EQ177     1           
    X575      EQ403     1              EQ472     -1          
    X575      EQ473     1                                    
    X576      EQ68      -1             EQ96      1           
    X576      EQ403     1              EQ472     -1          
    X576      EQ473     1                                    
    X577      EQ75      -1             EQ100     1           
    X577      EQ403     1              EQ472     -1          
    X577      EQ473     1                                    
    X578      EQ64      -1             EQ92      1           
    X578      EQ403     1              EQ472     -1          
    X578      EQ473     1                                    
    X579      EQ65      -1             EQ93      1           
    X579      EQ403     1              EQ472     -1          
    X579      EQ473     1                                    
    X580      EQ79      -1             EQ104     1           
    X580      EQ403     1              EQ472     -1          
    X580      EQ473     1                                    
    X581      EQ69      -1             EQ97      1           
    X581      EQ403     1              EQ472     -1          
    X581      EQ473     1                                    
    X582      EQ34      1              EQ72      -1          
    X582      EQ403     1              EQ472     -1          
    X582      EQ473     1                                    
    X583      EQ58      -1             EQ404     1           
    X583      EQ497     1                                    
    X584      EQ267     -1             EQ403     1           
    X584      EQ407     -1             EQ410     1           
    X585      EQ410     1              EQ470     1           
    X585      EQ741     -1             EQ756     -1          
    X586      EQ742     -1             EQ845     1           
    X587      EQ359     -1             EQ407     -1          
    X587      EQ414     1              EQ533     -1          
    X587      EQ547     1                                    
    X588      EQ201     1              EQ231     -1          
    X588      EQ363     1              EQ373     -1          
    X588      EQ403     1                                    
    X589      EQ129     1              EQ403     1           
    X589      EQ418     -1             EQ482     -1          
    X590      EQ73      -1             EQ403     1           
    X590      EQ466     1              EQ472     -1          
    X590      EQ473     1                                    
    X591      EQ419     -1             EQ478     1           
    X591      EQ686     1                                    
    X592      EQ403     3              EQ407     -1          
    X592      EQ419     1              EQ421     -1          
    X592      EQ472     -2             EQ473     2           
    X593      EQ407     -1             EQ420     -1          
    X593      EQ421     1              EQ538     1           
    X594      EQ217     1              EQ231     -1          
    X594      EQ419     -1             EQ422     1           
    X594      EQ547     1              EQ653     -1          
    X595      EQ201     1              EQ231     -1          
    X595      EQ403     1              EQ407     -1          
    X595      EQ419     1              EQ538     1           
    X595      EQ846     -1                                   
    X596      EQ743     -1             EQ846     1           
    X597      EQ407     -1             EQ423     1           
    X597      EQ478     4              EQ543     -4          
    X598      EQ180     1              EQ182     1           
    X598      EQ426     -1                                   
    X599      EQ112     -1             EQ113     1           
    X599      EQ201     1              EQ231     -1          
    X599      EQ403     1                                    
    X600      EQ30      1              EQ31      -1          
    X600      EQ403     3              EQ407     -1          
    X600      EQ472     -1             EQ473     1           
    X601      EQ403     2              EQ407     -1          
    X601      EQ424     1              EQ425     -1          
    X601      EQ472     -1             EQ473     1           
    X602      EQ31      -1             EQ355     1           
    X602      EQ403     2              EQ407     -1          
    X602      EQ493     1                                    
    X603      EQ29      1              EQ355     1           
    X603      EQ403     1              EQ407     -1          
    X603      EQ425     -1                                   
    X604      EQ178     1              EQ248     -1          
    X604      EQ262     1              EQ403     1           
    X604      EQ407     -1                                   
    X605      EQ119     -1             EQ181     1           
    X605      EQ581     1                                    
    X606      EQ123     -1             EQ542     1           
    X606      EQ581     1                                    
    X607      EQ120     -1             EQ181     1           
    X607      EQ581     1                                    
    X608      EQ20      1              EQ21      -1          
    X608      EQ217     1              EQ231     -1          
    X608      EQ403     1                                    
    X609      EQ28      1              EQ428     -1          
    X610      EQ228     1              EQ403     1           
    X610      EQ427     -1             EQ472     -1          
    X610      EQ473     1                                    
    X611      EQ182     -1             EQ187     1           
    X611      EQ269     1              EQ427     -1          
    X612      EQ201     1              EQ231     -1          
    X612      EQ403     1              EQ427     -1          
    X612      EQ535     1                                    
    X613      EQ208     1              EQ378     -1          
    X613      EQ420     1              EQ437     -1          
    X614      EQ403     1              EQ407     -1          
    X614      EQ430     -1             EQ472     -1          
    X614      EQ473     1              EQ694     1           
    X615      EQ744     -1             EQ847     1           
    X616      EQ430     -1             EQ438     1           
    X616      EQ547     1              EQ569     -1          
    X617      EQ738     -1             EQ842     1           
    X618      EQ208     1              EQ267     1           
    X618      EQ434     -1             EQ474     -1          
    X618      EQ475     1                                    
    X619      EQ403     1              EQ434     1           
    X619      EQ842     -1             EQ848     -1          
    X620      EQ745     -1             EQ848     1           
    X621      EQ387     1              EQ434     -1          
    X621      EQ613     1                                    
    X622      EQ206     1              EQ337     1           
    X622      EQ375     -1             EQ378     1           
    X622      EQ403     1              EQ566     -1          
    X623      EQ337     -1             EQ407     1           
    X623      EQ437     1                                    
    X624      EQ22      -1             EQ80      1           
    X624      EQ267     1              EQ403     -1          
    X624      EQ407     1                                    
    X625      EQ87      1              EQ208     -1          
    X625      EQ378     1              EQ435     -1          
    X626      EQ217     1              EQ231     -1          
    X626      EQ435     -1             EQ436     1           
    X626      EQ547     1              EQ654     -1          
    X627      EQ201     1              EQ231     -1          
    X627      EQ403     1              EQ407     -1          
    X627      EQ435     1              EQ538     1           
    X627      EQ849     -1                                   
    X628      EQ746     -1             EQ849     1           
    X629      EQ358     1              EQ407     -1          
    X629      EQ438     -1                                   
    X630      EQ403     1              EQ407     -1          
    X630      EQ438     -1             EQ472     -1          
    X630      EQ473     1              EQ695     1           
    X631      EQ407     -1             EQ430     1           
    X631      EQ440     -1             EQ592     1           
    X632      EQ323     1              EQ441     -1          
    X633      EQ403     -1             EQ406     -1          
    X633      EQ407     1              EQ441     1           
    X633      EQ472     1              EQ473     -1          
    X634      EQ52      -1             EQ54      1           
    X634      EQ403     1              EQ472     -1          
    X634      EQ473     1                                    
    X635      EQ32      1              EQ52      -1          
    X635      EQ407     1                                    
    X636      EQ32      -1             EQ53      1           
    X636      EQ407     -1                                   
    X637      EQ53      1              EQ86      -1          
    X637      EQ182     -1             EQ269     1           
    X637      EQ403     1              EQ407     -1          
    X638      EQ201     1              EQ231     -1          
    X638      EQ403     1              EQ407     -1          
    X638      EQ442     1              EQ538     1           
    X638      EQ852     -1                                   
    X639      EQ747     -1             EQ852     1           
    X640      EQ125     -1             EQ356     1           
    X640      EQ403     1              EQ407     -1          
    X641      EQ12      -1             EQ213     1           
    X641      EQ403     1              EQ474     -1          
    X641      EQ475     1                                    
    X642      EQ13      1              EQ19      -1          
    X642      EQ403     -1             EQ474     1           
    X642      EQ475     -1                                   
    X643      EQ153     1              EQ179     -1          
    X643      EQ191     1              EQ267     1           
    X643      EQ403     -1             EQ455     -1          
    X644      EQ182     -1             EQ191     1           
    X644      EQ267     1              EQ269     1           
    X644      EQ403     -1             EQ455     -1          
    X645      EQ177     -1             EQ182     2           
    X645      EQ269     -1                                   
    X646      EQ96      -1             EQ182     1           
    X646      EQ235     1              EQ269     -1          
    X647      EQ100     -1             EQ182     1           
    X647      EQ269     -1             EQ431     1           
    X648      EQ92      -1             EQ182     1           
    X648      EQ269     -1             EQ487     1           
    X649      EQ93      -1             EQ182     1           
    X649      EQ269     -1             EQ295     1           
    X650      EQ104     -1             EQ182     1           
    X650      EQ269     -1             EQ303     1           
    X651      EQ97      -1             EQ182     1           
    X651      EQ269     -1             EQ626     1           
    X652      EQ445     -1             EQ581     1           
    X652      EQ582     -1             EQ583     1           
    X653      EQ403     1              EQ445     1           
    X653      EQ842     -1             EQ854     -1          
    X654      EQ749     -1             EQ854     1           
    X655      EQ403     1              EQ444     -1          
    X655      EQ472     -1             EQ473     1           
    X655      EQ581     1                                    
    X656      EQ130     1              EQ208     -1          
    X656      EQ378     1              EQ446     -1          
    X657      EQ217     1              EQ231     -1          
    X657      EQ446     -1             EQ447     1           
    X657      EQ547     1              EQ655     -1          
    X658      EQ201     1              EQ231     -1          
    X658      EQ403     1              EQ407     -1          
    X658      EQ446     1              EQ538     1           
    X658      EQ855     -1                                   
    X659      EQ750     -1             EQ855     1           
    X660      EQ165     -1             EQ403     1           
    X660      EQ408     1              EQ482     -1          
    X660      EQ635     1                                    
    X661      EQ3       1              EQ267     1           
    X661      EQ403     -1             EQ450     -1          
    X662      EQ139     1              EQ267     1           
    X662      EQ403     1              EQ407     1           
    X662      EQ450     -1             EQ482     -1          
    X663      EQ449     1              EQ450     -1          
    X664      EQ217     1              EQ231     -1          
    X664      EQ450     -1             EQ451     1           
    X664      EQ547     1              EQ656     -1          
    X665      EQ201     1              EQ231     -1          
    X665      EQ403     1              EQ407     -1          
    X665      EQ450     1              EQ538     1           
    X665      EQ856     -1                                   
    X666      EQ751     -1             EQ856     1           
    X667      EQ453     1              EQ857     -1          
    X668      EQ752     -1             EQ857     1           
    X669      EQ117     1              EQ129     -1          
    X670      EQ34      -1             EQ182     1           
    X670      EQ269     -1             EQ544     1           
    X671      EQ179     1              EQ267     1           
    X671      EQ403     -1             EQ455     -1          
    X672      EQ360     1              EQ457     -1          
    X673      EQ403     1              EQ407     -1          
    X673      EQ457     1              EQ458     -1          
    X673      EQ477     1                                    
    X674      EQ182     -1             EQ269     1           
    X674      EQ387     -1             EQ403     1           
    X674      EQ407     -1             EQ454     1           
    X675      EQ403     2              EQ454     1           
    X675      EQ842     -2             EQ858     -1          
    X676      EQ753     -1             EQ858     1           
    X677      EQ81      1              EQ342     -1          
    X677      EQ343     1              EQ443     -1          
    X678      EQ81      -1             EQ85      1           
    X678      EQ201     1              EQ231     -1          
    X678      EQ403     1              EQ410     -1          
    X678      EQ538     1                                    
    X679      EQ38      1              EQ39      -1          
    X679      EQ407     1                                    
    X680      EQ464     -1             EQ581     1           
    X680      EQ613     1                                    
    X681      EQ39      1              EQ269     1           
    X681      EQ403     1              EQ407     -1          
    X681      EQ483     -1             EQ544     -1          
    X682      EQ153     -1             EQ269     1           
    X682      EQ455     1              EQ456     -1          
    X683      EQ403     1              EQ454     -1          
    X683      EQ472     -1             EQ473     1           
    X683      EQ483     1                                    
    X684      EQ454     -1             EQ483     1           
    X684      EQ582     -1             EQ583     1           
    X685      EQ267     1              EQ454     -1          
    X685      EQ472     -1             EQ473     1           
    X685      EQ581     1                                    
    X686      EQ41      -1             EQ403     -1          
    X686      EQ406     1              EQ407     1           
    X687      EQ41      1              EQ46      -1          
    X687      EQ265     1                                    
    X688      EQ40      -1             EQ114     1           
    X688      EQ281     -1             EQ403     -1          
    X688      EQ547     1                                    
    X689      EQ178     1              EQ187     -1          
    X689      EQ403     1              EQ409     -1          
    X689      EQ411     1                                    
    X690      EQ215     1              EQ231     -1          
    X690      EQ407     -1             EQ461     -1          
    X690      EQ538     1              EQ547     1           
    X691      EQ43      1              EQ258     1           
    X691      EQ407     -1             EQ461     -1          
    X691      EQ478     1                                    
    X692      EQ144     -1             EQ403     1           
    X692      EQ411     -1             EQ461     1           
    X692      EQ633     1                                    
    X693      EQ217     1              EQ231     -1          
    X693      EQ461     -1             EQ463     1           
    X693      EQ547     1              EQ657     -1          
    X694      EQ201     1              EQ231     -1          
    X694      EQ403     1              EQ407     -1          
    X694      EQ461     1              EQ538     1           
    X694      EQ859     -1                                   
    X695      EQ754     -1             EQ859     1           
    X696      EQ85      -1             EQ407     -1          
    X696      EQ426     1                                    
    X697      EQ38      -1             EQ407     -1          
    X697      EQ464     1                                    
    X698      EQ267     1              EQ269     -1          
    X698      EQ466     -1             EQ472     -1          
    X698      EQ473     1              EQ544     1           
    X699      EQ26      1              EQ86      -1          
    X699      EQ407     -1             EQ465     -1          
    X699      EQ633     1                                    
    X700      EQ201     1              EQ231     -1          
    X700      EQ403     1              EQ407     -1          
    X700      EQ467     1              EQ538     1           
    X700      EQ860     -1                                   
    X701      EQ755     -1             EQ860     1           
    X702      EQ258     -1             EQ403     4           
    X702      EQ407     -3             EQ467     1           
    X702      EQ474     -3             EQ475     3           
    X703      EQ1       1              EQ403     1           
    X703      EQ407     -1             EQ462     -1          
    X704      EQ462     1              EQ465     -1          
    X704      EQ474     -1             EQ475     1           
    X705      EQ144     1              EQ403     -2          
    X705      EQ465     -1             EQ474     1           
    X705      EQ475     -1                                   
    X706      EQ145     -1             EQ411     -1          
    X706      EQ461     1              EQ634     1           
    X707      EQ308     1              EQ403     -1          
    X707      EQ452     -1             EQ472     1           
    X707      EQ473     -1             EQ482     -1          
    X708      EQ116     1              EQ162     -1          
    X708      EQ403     -1                                   
    X709      EQ249     -1             EQ403     -1          
    X709      EQ468     1                                    
    X710      EQ146     1              EQ468     -1          
    X711      EQ82      1              EQ403     -1          
    X711      EQ407     1              EQ453     -1          
    X711      EQ472     1              EQ473     -1          
    X711      EQ482     -1                                   
    X712      EQ149     1              EQ471     -1          
    X712      EQ482     -0.5                                 
    X713      EQ403     1              EQ471     1           
    X713      EQ842     -1             EQ862     -1          
    X714      EQ757     -1             EQ862     1           
    X715      EQ403     -4             EQ472     1           
    X715      EQ473     -1             EQ582     -1          
    X715      EQ583     1              EQ842     3           
    X716      EQ403     -1             EQ472     1           
    X716      EQ473     -1             EQ582     -1          
    X716      EQ583     1                                    
    X717      EQ201     1              EQ231     -1          
    X717      EQ403     1              EQ472     -1          
    X717      EQ474     1                                    
    X718      EQ403     -1             EQ474     1           
    X718      EQ475     -1             EQ582     -1          
    X718      EQ583     1                                    
    X719      EQ217     1              EQ231     -1          
    X719      EQ326     -1             EQ375     -1          
    X719      EQ378     1              EQ403     1           
    X719      EQ407     -1             EQ472     1           
    X719      EQ547     1                                    
    X720      EQ201     1              EQ231     -1          
    X720      EQ407     -1             EQ471     -1          
    X720      EQ480     1              EQ538     1           
    X720      EQ547     1              EQ569     -1          
    X721      EQ403     3              EQ470     -2          
    X721      EQ842     -3             EQ861     2           
    X722      EQ403     2              EQ470     -1          
    X722      EQ842     -2             EQ861     1           
    X723      EQ756     -1             EQ861     1           
    X724      EQ201     1              EQ231     -1          
    X724      EQ371     -1             EQ401     1           
    X725      EQ201     1              EQ231     -1          
    X725      EQ677     -1             EQ689     1           
    X726      EQ201     1              EQ231     -1          
    X726      EQ252     -1             EQ281     1           
    X727      EQ201     1              EQ231     -1          
    X727      EQ329     -1             EQ331     1           
    X728      EQ201     1              EQ231     -1          
    X728      EQ304     -1             EQ307     1           
    X729      EQ201     1              EQ231     -1          
    X729      EQ332     -1             EQ334     1           
    X730      EQ201     1              EQ231     -1          
    X730      EQ297     -1             EQ299     1           
    X731      EQ201     1              EQ231     -1          
    X731      EQ289     -1             EQ291     1           
    X732      EQ355     1              EQ403     1           
    X732      EQ407     -1             EQ458     1           
    X732      EQ476     -1                                   
    X733      EQ166     -1             EQ355     1           
    X733      EQ378     1              EQ403     2           
    X733      EQ407     -1                                   
    X734      EQ758     -1             EQ863     1           
    X735      EQ478     1              EQ863     -1          
    X736      EQ759     -1             EQ864     1           
    X737      EQ201     1              EQ231     -1          
    X737      EQ403     1              EQ407     -1          
    X737      EQ479     1              EQ538     1           
    X737      EQ864     -1                                   
    X738      EQ231     -1             EQ326     1           
    X738      EQ403     -1             EQ480     -1          
    X738      EQ547     1                                    
    X739      EQ147     1              EQ320     -1          
    X739      EQ403     1              EQ471     1           
    X739      EQ480     -1                                   
    X740      EQ267     1              EQ403     -2          
    X740      EQ480     1              EQ547     1           
    X740      EQ569     -1             EQ585     -1          
    X741      EQ403     1              EQ481     1           
    X741      EQ842     -1             EQ865     -1          
    X742      EQ760     -1             EQ865     1           
    X743      EQ407     -1             EQ538     1           
    X743      EQ695     -1             EQ696     1           
    X744      EQ844     -1             EQ888     1           
    X744      EQ910     -1             EQ911     1           
    X745      EQ407     -1             EQ438     -1          
    X745      EQ440     1              EQ538     1           
    X746      EQ844     -1             EQ850     -1          
    X746      EQ851     1              EQ888     1           
    X747      EQ407     -1             EQ538     1           
    X747      EQ681     -1             EQ688     1           
    X748      EQ844     -1             EQ888     1           
    X748      EQ905     -1             EQ906     1           
    X749      EQ298     -1             EQ300     1           
    X749      EQ407     -1             EQ538     1           
    X750      EQ822     -1             EQ823     1           
    X750      EQ844     -1             EQ888     1           
    X751      EQ265     -1             EQ286     1           
    X751      EQ407     -1             EQ538     1           
    X752      EQ812     -1             EQ818     1           
    X752      EQ844     -1             EQ888     1           
    X753      EQ330     -1             EQ407     -1          
    X753      EQ538     1              EQ638     1           
    X754      EQ827     -1             EQ844     -1          
    X754      EQ888     1              EQ900     1           
    X755      EQ287     1              EQ290     -1          
    X755      EQ407     -1             EQ538     1           
    X756      EQ819     1              EQ820     -1          
    X756      EQ844     -1             EQ888     1           
    X757      EQ197     1              EQ217     -1          
    X757      EQ407     -1             EQ538     1           
    X758      EQ796     1              EQ799     -1          
    X758      EQ844     -1             EQ888     1           
    X759      EQ305     -1             EQ306     1           
    X759      EQ407     -1             EQ538     1           
    X760      EQ825     -1             EQ826     1           
    X760      EQ844     -1             EQ888     1           
    X761      EQ396     -1             EQ398     1           
    X761      EQ407     -1             EQ538     1           
    X762      EQ840     -1             EQ841     1           
    X762      EQ844     -1             EQ888     1           
    X763      EQ403     -5             EQ407     2           
    X763      EQ472     3              EQ473     -3          
    X763      EQ478     1              EQ481     -1          
    X764      EQ761     -1             EQ866     1           
    X765      EQ482     1              EQ866     -1          
    X766      EQ267     1              EQ403     -1          
    X766      EQ483     -1             EQ581     1           
    X767      EQ267     1              EQ370     1           
    X767      EQ403     -1             EQ499     -1          
    X768      EQ43      -1             EQ267     1           
    X768      EQ269     -1             EQ472     -1          
    X768      EQ473     1              EQ544     1           
    X769      EQ483     1              EQ486     -1          
    X769      EQ581     1                                    
    X770      EQ247     -1             EQ264     1           
    X770      EQ403     1              EQ494     -1          
    X770      EQ538     1                                    
    X771      EQ177     1              EQ180     -1          
    X771      EQ613     1              EQ614     -1          
    X772      EQ762     -1             EQ867     1           
    X773      EQ208     1              EQ378     -1          
    X773      EQ490     -1             EQ537     1           
    X774      EQ130     -1             EQ267     1           
    X774      EQ269     -1             EQ443     1           
    X774      EQ472     -1             EQ473     1           
    X775      EQ86      -1             EQ267     1           
    X775      EQ269     -1             EQ433     1           
    X775      EQ472     -1             EQ473     1           
    X776      EQ37      1              EQ87      -1          
    X776      EQ267     1              EQ269     -1          
    X776      EQ472     -1             EQ473     1           
    X777      EQ407     -1             EQ486     1           
    X777      EQ491     -1                                   
    X778      EQ491     1              EQ492     -1          
    X779      EQ54      -1             EQ130     1           
    X779      EQ267     1              EQ403     -1          
    X780      EQ267     1              EQ403     -1          
    X780      EQ496     -1             EQ681     1           
    X781      EQ119     1              EQ407     -1          
    X781      EQ493     -1                                   
    X782      EQ29      -1             EQ123     1           
    X782      EQ407     -1                                   
    X783      EQ120     1              EQ262     -1          
    X783      EQ407     -1                                   
    X784      EQ478     1              EQ494     -1          
    X784      EQ567     1                                    
    X785      EQ267     1              EQ403     -1          
    X785      EQ494     -1             EQ577     1           
    X786      EQ186     1              EQ189     -1          
    X786      EQ378     -1             EQ494     1           
    X787      EQ201     1              EQ231     -1          
    X787      EQ403     1              EQ407     -1          
    X787      EQ494     1              EQ538     1           
    X787      EQ868     -1                                   
    X788      EQ763     -1             EQ868     1           
    X789      EQ495     1              EQ496     -1          
    X789      EQ547     -1             EQ569     1           
    X790      EQ45      -1             EQ174     1           
    X790      EQ267     1              EQ403     -1          
    X790      EQ448     -1                                   
    X791      EQ105     1              EQ146     -1          
    X791      EQ403     1              EQ407     -1          
    X792      EQ383     1              EQ403     2           
    X792      EQ407     -1             EQ472     -1          
    X792      EQ473     1              EQ500     -1          
    X793      EQ502     1              EQ869     -1          
    X794      EQ764     -1             EQ869     1           
    X795      EQ10      -1             EQ378     1           
    X795      EQ403     1              EQ407     -2          
    X795      EQ472     -1             EQ473     1           
    X796      EQ10      -1             EQ403     -2          
    X796      EQ474     1              EQ475     -1          
    X796      EQ567     1                                    
    X797      EQ111     1              EQ261     -1          
    X797      EQ407     1              EQ478     -1          
    X798      EQ217     1              EQ231     -1          
    X798      EQ269     -1             EQ508     -1          
    X798      EQ531     1              EQ547     1           
    X799      EQ765     -1             EQ870     1           
    X800      EQ470     1              EQ508     1           
    X800      EQ861     -1             EQ870     -1          
    X801      EQ209     -1             EQ217     1           
    X801      EQ231     -1             EQ403     1           
    X801      EQ511     -1             EQ540     1           
    X801      EQ547     1                                    
    X802      EQ671     1              EQ676     -1          
    X802      EQ680     -1             EQ681     1           
    X803      EQ201     1              EQ231     -1          
    X803      EQ403     1              EQ410     -1          
    X803      EQ483     1              EQ538     1           
    X803      EQ581     -1                                   
    X804      EQ204     1              EQ215     -1          
    X804      EQ403     1              EQ560     -1          
    X804      EQ561     1                                    
    X805      EQ204     1              EQ215     -1          
    X805      EQ403     3              EQ559     -1          
    X805      EQ560     1                                    
    X806      EQ204     1              EQ215     -1          
    X806      EQ328     -1             EQ403     1           
    X806      EQ558     1                                    
    X807      EQ403     -1             EQ474     1           
    X807      EQ475     -1             EQ562     -1          
    X807      EQ563     1                                    
    X808      EQ204     2              EQ215     -2          
    X808      EQ267     1              EQ403     2           
    X808      EQ563     -1             EQ564     1           
    X809      EQ403     -1             EQ417     1           
    X809      EQ564     -1                                   
    X810      EQ49      -1             EQ162     1           
    X810      EQ403     2              EQ482     -1          
    X811      EQ182     1              EQ267     1           
    X811      EQ269     -1             EQ472     -1          
    X811      EQ473     1              EQ581     -1          
    X812      EQ196     1              EQ267     1           
    X812      EQ403     -1             EQ448     -1          
    X812      EQ581     -1                                   
    X813      EQ182     1              EQ196     -1          
    X813      EQ269     -1             EQ313     1           
    X814      EQ313     -1             EQ403     1           
    X814      EQ448     1              EQ472     -1          
    X814      EQ473     1                                    
    X815      EQ408     1              EQ482     -1          
    X815      EQ512     -1             EQ579     1           
    X816      EQ267     1              EQ335     -1          
    X816      EQ403     1              EQ407     2           
    X816      EQ472     -1             EQ473     1           
    X816      EQ512     1              EQ537     -1          
    X816      EQ538     1                                    
    X817      EQ407     -1             EQ512     -1          
    X817      EQ538     1              EQ580     1           
    X818      EQ201     1              EQ231     -1          
    X818      EQ403     1              EQ407     -1          
    X818      EQ513     -1             EQ538     1           
    X818      EQ871     1                                    
    X819      EQ201     1              EQ231     -1          
    X819      EQ403     1              EQ407     -1          
    X819      EQ514     -1             EQ538     1           
    X819      EQ872     1                                    
    X820      EQ201     1              EQ231     -1          
    X820      EQ403     1              EQ407     -1          
    X820      EQ515     -1             EQ538     1           
    X820      EQ873     1                                    
    X821      EQ201     1              EQ231     -1          
    X821      EQ403     1              EQ407     -1          
    X821      EQ516     -1             EQ538     1           
    X821      EQ874     1                                    
    X822      EQ201     1              EQ231     -1          
    X822      EQ403     1              EQ407     -1          
    X822      EQ517     -1             EQ538     1           
    X822      EQ875     1                                    
    X823      EQ766     -1             EQ876     1           
    X824      EQ132     -1             EQ403     1           
    X824      EQ472     -1             EQ473     1           
    X824      EQ490     1                                    
    X825      EQ366     -1             EQ367     1           
    X826      EQ106     -1             EQ107     1           
    X826      EQ403     1              EQ472     -1          
    X826      EQ473     1                                    
    X827      EQ341     1              EQ363     -1          
    X828      EQ2       1              EQ106     -1          
    X828      EQ201     1              EQ231     -1          
    X829      EQ151     1              EQ152     -1          
    X829      EQ403     1              EQ407     -1          
    X830      EQ150     -1             EQ151     1           
    X830      EQ403     2              EQ474     1           
    X830      EQ475     -1                                   
    X831      EQ47      -1             EQ106     1           
    X832      EQ361     -1             EQ363     1           
    X833      EQ201     1              EQ231     -1          
    X833      EQ403     1              EQ407     -1          
    X833      EQ526     -1             EQ538     1           
    X833      EQ882     1                                    
    X834      EQ201     1              EQ231     -1          
    X834      EQ403     1              EQ407     -1          
    X834      EQ527     -1             EQ538     1           
    X834      EQ883     1                                    
    X835      EQ201     1              EQ231     -1          
    X835      EQ403     1              EQ407     -1          
    X835      EQ528     -1             EQ538     1           
    X835      EQ884     1                                    
    X836      EQ201     1              EQ231     -1          
    X836      EQ403     1              EQ407     -1          
    X836      EQ529     -1             EQ538     1           
    X836      EQ885     1                                    
    X837      EQ201     1              EQ231     -1          
    X837      EQ403     1              EQ407     -1          
    X837      EQ530     -1             EQ538     1           
    X837      EQ886     1                                    
    X838      EQ407     -1             EQ521     1           
    X838      EQ526     -1             EQ538     1           
    X839      EQ844     -1             EQ877     1           
    X839      EQ882     -1             EQ888     1           
    X840      EQ407     -1             EQ522     1           
    X840      EQ527     -1             EQ538     1           
    X841      EQ844     -1             EQ878     1           
    X841      EQ883     -1             EQ888     1           
    X842      EQ407     -1             EQ523     1           
    X842      EQ528     -1             EQ538     1           
    X843      EQ844     -1             EQ879     1           
    X843      EQ884     -1             EQ888     1           
    X844      EQ407     -1             EQ524     1           
    X844      EQ529     -1             EQ538     1           
    X845      EQ844     -1             EQ880     1           
    X845      EQ885     -1             EQ888     1           
    X846      EQ407     -1             EQ525     1           
    X846      EQ530     -1             EQ538     1           
    X847      EQ844     -1             EQ881     1           
    X847      EQ886     -1             EQ888     1           
    X848      EQ253     -1             EQ265     1           
    X848      EQ393     -1             EQ403     1           
    X848      EQ526     1                                    
    X849      EQ255     -1             EQ265     1           
    X849      EQ393     -1             EQ403     1           
    X849      EQ527     1                                    
    X850      EQ254     -1             EQ265     1           
    X850      EQ393     -1             EQ403     1           
    X850      EQ528     1                                    
    X851      EQ257     -1             EQ265     1           
    X851      EQ393     -1             EQ403     1           
    X851      EQ529     1                                    
    X852      EQ256     -1             EQ265     1           
    X852      EQ393     -1             EQ403     1           
    X852      EQ530     1                                    
    X853      EQ155     -1             EQ156     -1          
    X853      EQ160     -1             EQ161     -1          
    X853      EQ459     1                                    
    X854      EQ27      1              EQ404     -1          
    X854      EQ482     -1             EQ531     -1          
    X855      EQ155     1              EQ169     -1          
    X855      EQ269     1              EQ460     -1          
    X856      EQ156     1              EQ170     -1          
    X856      EQ269     1              EQ460     -1          
    X857      EQ157     1              EQ167     -1          
    X857      EQ269     1              EQ460     -1          
    X858      EQ158     1              EQ172     -1          
    X858      EQ269     1              EQ460     -1          
    X859      EQ159     1              EQ168     -1          
    X859      EQ269     1              EQ460     -1          
    X860      EQ160     1              EQ171     -1          
    X860      EQ269     1              EQ460     -1          
    X861      EQ161     1              EQ173     -1          
    X861      EQ269     1              EQ460     -1          
    X862      EQ208     -1             EQ378     1           
    X862      EQ532     -1             EQ536     1           
    X863      EQ311     1              EQ407     1           
    X863      EQ482     -1             EQ532     -1          
    X863      EQ630     -1             EQ666     1           
    X864      EQ217     1              EQ231     -1          
    X864      EQ532     -1             EQ534     1           
    X864      EQ547     1              EQ658     -1          
    X865      EQ403     1              EQ532     1           
    X865      EQ842     -1             EQ887     -1          
    X866      EQ767     -1             EQ887     1           
    X867      EQ403     1              EQ538     1           
    X867      EQ842     -1             EQ888     -1          
    X868      EQ768     -1             EQ888     1           
    X869      EQ201     1              EQ231     -1          
    X869      EQ403     1              EQ407     -1          
    X869      EQ538     2              EQ888     -1          
    X870      EQ135     1              EQ140     -1          
    X870      EQ407     -1             EQ538     1           
    X871      EQ35      1              EQ113     -1          
    X871      EQ201     1              EQ231     -1          
    X872      EQ133     1              EQ201     1           
    X872      EQ231     -1             EQ403     1           
    X872      EQ540     -1                                   
    X873      EQ403     1              EQ407     -1          
    X873      EQ538     2              EQ547     -1          
    X874      EQ470     1              EQ541     1           
    X874      EQ861     -1             EQ889     -1          
    X875      EQ769     -1             EQ889     1           
    X876      EQ138     -2             EQ403     1           
    X876      EQ407     2              EQ543     1           
    X877      EQ267     -1             EQ403     1           
    X877      EQ407     -1             EQ483     1           
    X877      EQ519     -1             EQ538     1           
    X878      EQ134     -1             EQ267     1           
    X878      EQ403     -1             EQ510     1           
    X879      EQ371     1              EQ407     -1          
    X879      EQ545     -1             EQ547     1           
    X880      EQ201     1              EQ231     -1          
    X880      EQ547     -1             EQ550     1           
    X881      EQ587     -1             EQ588     1           
    X882      EQ133     -1             EQ134     1           
    X882      EQ252     1              EQ281     -1          
    X882      EQ282     -1             EQ403     1           
    X882      EQ538     1                                    
    X883      EQ51      1              EQ267     1           
    X883      EQ472     -1             EQ473     1           
    X883      EQ546     -1                                   
    X884      EQ267     1              EQ403     -1          
    X884      EQ407     1              EQ536     1           
    X884      EQ546     -1                                   
    X885      EQ407     3              EQ482     -1.5        
    X885      EQ548     1              EQ549     -1          
    X886      EQ142     1              EQ403     1           
    X886      EQ407     -2             EQ472     -1          
    X886      EQ473     1              EQ551     -1          
    X887      EQ217     1              EQ231     -1          
    X887      EQ403     2              EQ407     -1          
    X887      EQ519     1              EQ538     1           
    X887      EQ581     -1                                   
    X888      EQ403     1              EQ520     1           
    X888      EQ668     -1             EQ675     1           
    X889      EQ201     1              EQ231     -1          
    X889      EQ368     1              EQ388     -1          
    X889      EQ403     1              EQ538     1           
    X889      EQ554     -1                                   
    X890      EQ22      1              EQ555     -1          
    X891      EQ201     1              EQ207     1           
    X891      EQ231     -1             EQ357     -1          
    X891      EQ403     2              EQ538     1           
    X892      EQ407     -1             EQ556     -1          
    X892      EQ565     1                                    
    X893      EQ15      1              EQ137     -1          
    X893      EQ201     1              EQ227     -1          
    X893      EQ231     -1             EQ403     1           
    X893      EQ538     1                                    
    X894      EQ403     1              EQ407     -1          
    X894      EQ547     1              EQ556     1           
    X894      EQ557     -1                                   
    X895      EQ201     1              EQ231     -1          
    X895      EQ349     -1             EQ357     1           
    X895      EQ375     -1             EQ378     1           
    X895      EQ403     1              EQ407     -1          
    X895      EQ538     1                                    
    X896      EQ565     -1             EQ566     1           
    X897      EQ10      1              EQ342     -1          
    X897      EQ343     1              EQ403     1           
    X897      EQ567     -1                                   
    X898      EQ217     1              EQ231     -1          
    X898      EQ547     1              EQ567     -1          
    X898      EQ568     1              EQ659     -1          
    X899      EQ201     1              EQ231     -1          
    X899      EQ403     1              EQ407     -1          
    X899      EQ538     1              EQ567     1           
    X899      EQ890     -1                                   
    X900      EQ403     1              EQ567     1           
    X900      EQ842     -1             EQ890     -1          
    X901      EQ470     1              EQ567     1           
    X901      EQ861     -1             EQ890     -1          
    X902      EQ770     -1             EQ890     1           
    X903      EQ217     1              EQ231     -1          
    X903      EQ403     1              EQ569     1           
    X903      EQ588     -1                                   
    X904      EQ108     1              EQ519     -1          
    X904      EQ538     1              EQ606     -1          
    X905      EQ267     1              EQ403     -1          
    X905      EQ513     1              EQ570     -1          
    X906      EQ267     1              EQ403     -1          
    X906      EQ514     1              EQ571     -1          
    X907      EQ267     1              EQ403     -1          
    X907      EQ515     1              EQ572     -1          
    X908      EQ267     1              EQ403     -1          
    X908      EQ516     1              EQ573     -1          
    X909      EQ267     1              EQ403     -1          
    X909      EQ517     1              EQ574     -1          
    X910      EQ107     -1             EQ208     1           
    X910      EQ378     -1             EQ575     1           
    X911      EQ407     -1             EQ538     1           
    X911      EQ575     -1             EQ602     1           
    X912      EQ253     -1             EQ265     1           
    X912      EQ403     1              EQ570     1           
    X912      EQ602     -1                                   
    X913      EQ255     -1             EQ265     1           
    X913      EQ403     1              EQ571     1           
    X913      EQ602     -1                                   
    X914      EQ254     -1             EQ265     1           
    X914      EQ403     1              EQ572     1           
    X914      EQ602     -1                                   
    X915      EQ257     -1             EQ265     1           
    X915      EQ403     1              EQ573     1           
    X915      EQ602     -1                                   
    X916      EQ256     -1             EQ265     1           
    X916      EQ403     1              EQ574     1           
    X916      EQ602     -1                                   
    X917      EQ182     -1             EQ192     1           
    X917      EQ269     1              EQ538     -1          
    X918      EQ309     1              EQ403     -1          
    X918      EQ472     1              EQ473     -1          
    X918      EQ482     -1             EQ501     -1          
    X919      EQ231     -1             EQ327     1           
    X919      EQ403     -1             EQ510     -1          
    X919      EQ547     1                                    
    X920      EQ201     1              EQ231     -1          
    X920      EQ403     1              EQ407     -1          
    X920      EQ538     1              EQ577     1           
    X920      EQ891     -1                                   
    X921      EQ771     -1             EQ891     1           
    X922      EQ126     1              EQ403     -1          
    X922      EQ407     1              EQ472     1           
    X922      EQ473     -1             EQ482     -1          
    X922      EQ502     -1                                   
    X923      EQ407     -1             EQ538     1           
    X923      EQ578     1              EQ579     -1          
    X924      EQ201     -1             EQ231     1           
    X924      EQ403     -1             EQ519     -1          
    X924      EQ581     1                                    
    X925      EQ538     -1             EQ587     1           
    X925      EQ684     1              EQ688     -1          
    X926      EQ318     -1             EQ403     1           
    X926      EQ476     1              EQ482     -1          
    X927      EQ55      1              EQ552     -1          
    X927      EQ553     1              EQ584     -1          
    X928      EQ772     -1             EQ892     1           
    X929      EQ403     1              EQ584     1           
    X929      EQ842     -1             EQ892     -1          
    X930      EQ310     -1             EQ407     2           
    X930      EQ432     -1             EQ538     1           
    X930      EQ585     1                                    
    X931      EQ403     -1             EQ482     -0.5        
    X931      EQ558     -1             EQ559     1           
    X932      EQ178     1              EQ204     1           
    X932      EQ215     -1             EQ403     1           
    X932      EQ407     -1             EQ561     -1          
    X932      EQ562     1                                    
    X933      EQ201     2              EQ231     -2          
    X933      EQ375     -2             EQ378     2           
    X933      EQ403     1              EQ407     -1          
    X933      EQ416     1              EQ417     -1          
    X933      EQ547     1                                    
    X934      EQ201     1              EQ231     -1          
    X934      EQ569     1              EQ586     -1          
    X935      EQ201     1              EQ231     -1          
    X935      EQ403     1              EQ586     1           
    X935      EQ587     -1                                   
    X936      EQ201     1              EQ231     -1          
    X936      EQ352     1              EQ403     1           
    X936      EQ593     -1                                   
    X937      EQ135     -1             EQ292     -1          
    X937      EQ322     1              EQ407     2           
    X937      EQ538     1                                    
    X938      EQ135     1              EQ322     -2          
    X938      EQ593     1                                    
    X939      EQ201     1              EQ231     -1          
    X939      EQ403     1              EQ588     1           
    X939      EQ592     -1                                   
    X940      EQ92      -1             EQ169     1           
    X940      EQ403     -1             EQ474     1           
    X940      EQ475     -1                                   
    X941      EQ93      -1             EQ170     1           
    X941      EQ403     -1             EQ474     1           
    X941      EQ475     -1                                   
    X942      EQ104     -1             EQ172     1           
    X942      EQ403     -1             EQ474     1           
    X942      EQ475     -1                                   
    X943      EQ96      -1             EQ171     1           
    X943      EQ403     -1             EQ474     1           
    X943      EQ475     -1                                   
    X944      EQ100     -1             EQ173     1           
    X944      EQ403     -1             EQ474     1           
    X944      EQ475     -1                                   
    X945      EQ201     1              EQ231     -1          
    X945      EQ403     1              EQ407     -1          
    X945      EQ538     1              EQ592     1           
    X945      EQ893     -1                                   
    X946      EQ773     -1             EQ893     1           
    X947      EQ403     1              EQ472     1           
    X947      EQ473     -1             EQ589     -1          
    X947      EQ590     1                                    
    X948      EQ201     -1             EQ289     1           
    X948      EQ407     1              EQ644     1           
    X948      EQ645     -1                                   
    X949      EQ304     1              EQ371     -1          
    X949      EQ407     1              EQ644     1           
    X949      EQ645     -1                                   
    X950      EQ252     -1             EQ297     1           
    X950      EQ407     1              EQ644     1           
    X950      EQ645     -1                                   
    X951      EQ332     1              EQ407     1           
    X951      EQ644     1              EQ645     -1          
    X951      EQ677     -1                                   
    X952      EQ594     -1             EQ697     1           
    X953      EQ588     -1             EQ594     1           
    X954      EQ147     -1             EQ407     -1          
    X954      EQ538     1              EQ591     1           
    X955      EQ267     1              EQ403     -2          
    X955      EQ407     -2             EQ478     2           
    X955      EQ612     -1             EQ618     1           
    X956      EQ222     1              EQ231     -1          
    X956      EQ403     -1             EQ547     1           
    X956      EQ610     -1                                   
    X957      EQ344     1              EQ388     1           
    X957      EQ407     -1             EQ408     1           
    X957      EQ482     -1             EQ596     -1          
    X958      EQ282     1              EQ399     1           
    X958      EQ400     -2             EQ403     1           
    X958      EQ598     -1             EQ609     1           
    X959      EQ17      1              EQ407     -1          
    X959      EQ607     -1             EQ613     1           
    X960      EQ208     -1             EQ378     1           
    X960      EQ607     -1             EQ608     1           
    X961      EQ231     -1             EQ407     -1          
    X961      EQ538     1              EQ547     1           
    X961      EQ600     1              EQ601     -1          
    X962      EQ182     -1             EQ190     1           
    X962      EQ269     1              EQ602     -1          
    X963      EQ478     1              EQ581     1           
    X963      EQ602     -1                                   
    X964      EQ217     1              EQ231     -1          
    X964      EQ547     1              EQ602     -1          
    X964      EQ603     1              EQ660     -1          
    X965      EQ403     1              EQ477     1           
    X965      EQ581     1              EQ602     -1          
    X966      EQ403     1              EQ602     1           
    X966      EQ842     -1             EQ894     -1          
    X967      EQ774     -1             EQ894     1           
    X968      EQ378     1              EQ407     -1          
    X968      EQ613     1              EQ615     -1          
    X969      EQ403     2              EQ407     -1          
    X969      EQ472     -1             EQ473     1           
    X969      EQ615     1              EQ616     -1          
    X970      EQ328     -1             EQ403     1           
    X970      EQ472     -1             EQ473     1           
    X970      EQ597     1                                    
    X971      EQ346     -1             EQ403     3           
    X971      EQ597     -1             EQ604     1           
    X972      EQ56      -1             EQ403     -1          
    X972      EQ474     1              EQ475     -1          
    X972      EQ605     1                                    
    X973      EQ201     1              EQ231     -1          
    X973      EQ403     1              EQ605     -1          
    X973      EQ606     1                                    
    X974      EQ282     -1             EQ283     1           
    X974      EQ403     1              EQ613     1           
    X974      EQ617     -1                                   
    X975      EQ403     1              EQ409     -1          
    X975      EQ411     1              EQ613     1           
    X975      EQ617     -1                                   
    X976      EQ43      1              EQ403     1           
    X976      EQ407     -1             EQ478     1           
    X976      EQ613     1              EQ617     -1          
    X977      EQ470     2              EQ610     1           
    X977      EQ756     -2             EQ775     -1          
    X978      EQ775     -1             EQ895     1           
    X979      EQ208     -1             EQ378     1           
    X979      EQ616     1              EQ618     -1          
    X980      EQ201     1              EQ231     -1          
    X980      EQ403     1              EQ407     -1          
    X980      EQ538     1              EQ611     1           
    X980      EQ896     -1                                   
    X981      EQ776     -1             EQ896     1           
    X982      EQ143     1              EQ216     -1          
    X982      EQ403     1              EQ577     -1          
    X982      EQ611     1                                    
    X983      EQ403     2              EQ407     -1          
    X983      EQ474     -1             EQ475     1           
    X983      EQ613     1              EQ619     -1          
    X984      EQ403     2              EQ613     1           
    X984      EQ842     -2             EQ897     -1          
    X985      EQ777     -1             EQ897     1           
    X986      EQ342     -1             EQ343     1           
    X986      EQ360     1              EQ613     -1          
    X987      EQ360     1              EQ582     -1          
    X987      EQ583     1              EQ613     -1          
    X988      EQ201     1              EQ231     -1          
    X988      EQ269     -1             EQ538     1           
    X988      EQ613     -1             EQ614     1           
    X989      EQ403     -5             EQ407     3           
    X989      EQ409     1              EQ474     3           
    X989      EQ475     -3             EQ609     -1          
    X990      EQ201     1              EQ231     -1          
    X990      EQ403     1              EQ407     -1          
    X990      EQ538     1              EQ610     1           
    X990      EQ895     -1                                   
    X991      EQ251     1              EQ629     -1          
    X992      EQ705     -1             EQ793     1           
    X993      EQ175     1              EQ793     -1          
    X994      EQ336     1              EQ341     1           
    X994      EQ362     -1             EQ595     -1          
    X995      EQ176     1              EQ208     -1          
    X995      EQ267     1              EQ403     1           
    X995      EQ482     -1             EQ609     1           
    X995      EQ613     1              EQ624     -1          
    X996      EQ201     1              EQ231     -1          
    X996      EQ403     1              EQ407     -1          
    X996      EQ538     1              EQ624     1           
    X996      EQ898     -1                                   
    X997      EQ778     -1             EQ898     1           
    X998      EQ472     -1             EQ473     1           
    X998      EQ474     1              EQ475     -1          
    X999      EQ269     1              EQ407     -1          
    X999      EQ608     1              EQ614     -1          
    X999      EQ631     -1                                   
    X1000     EQ181     1              EQ230     -1          
    X1000     EQ388     1                                    
    X1001     EQ181     1              EQ388     1           
    X1001     EQ636     -1                                   
    X1002     EQ43      1              EQ478     1           
    X1002     EQ636     -1                                   
    X1003     EQ407     -1             EQ535     -1          
    X1003     EQ538     1              EQ636     1           
    X1004     EQ217     1              EQ231     -1          
    X1004     EQ547     1              EQ636     -1          
    X1004     EQ637     1              EQ661     -1          
    X1005     EQ201     1              EQ231     -1          
    X1005     EQ403     1              EQ407     -1          
    X1005     EQ538     1              EQ636     1           
    X1005     EQ899     -1                                   
    X1006     EQ779     -1             EQ899     1           
    X1007     EQ362     1              EQ588     -1          
    X1007     EQ595     1              EQ697     -1          
    X1008     EQ336     -1             EQ341     1           
    X1008     EQ362     1              EQ697     -1          
    X1009     EQ312     1              EQ330     1           
    X1009     EQ333     -1             EQ465     -1          
    X1010     EQ780     -1             EQ901     1           
    X1011     EQ642     1              EQ901     -1          
    X1012     EQ310     -1             EQ362     1           
    X1013     EQ403     -1             EQ474     1           
    X1013     EQ475     -1             EQ644     -1          
    X1013     EQ645     1                                    
    X1014     EQ80      -1             EQ362     1           
    X1014     EQ407     1              EQ602     -1          
    X1014     EQ664     1                                    
    X1015     EQ407     1              EQ439     -1          
    X1015     EQ602     -1             EQ664     1           
    X1016     EQ80      -1             EQ362     1           
    X1016     EQ439     1                                    
    X1017     EQ28      -1             EQ392     1           
    X1017     EQ403     -1             EQ472     1           
    X1017     EQ473     -1                                   
    X1018     EQ201     1              EQ231     -1          
    X1018     EQ403     1              EQ407     -1          
    X1018     EQ538     1              EQ665     1           
    X1018     EQ902     -1                                   
    X1019     EQ781     -1             EQ902     1           
    X1020     EQ782     -1             EQ903     1           
    X1021     EQ51      1              EQ208     -1          
    X1021     EQ378     1              EQ666     -1          
    X1022     EQ217     1              EQ231     -1          
    X1022     EQ547     1              EQ662     -1          
    X1022     EQ666     -1             EQ667     1           
    X1023     EQ403     1              EQ666     1           
    X1023     EQ842     -1             EQ904     -1          
    X1024     EQ783     -1             EQ904     1           
    X1025     EQ18      -1             EQ201     1           
    X1025     EQ231     -1             EQ403     1           
    X1025     EQ538     1              EQ673     -1          
    X1025     EQ679     1                                    
    X1026     EQ519     -1             EQ538     1           
    X1026     EQ669     1              EQ670     -1          
    X1027     EQ185     -1             EQ403     -1          
    X1027     EQ547     1              EQ670     1           
    X1027     EQ689     -1                                   
    X1028     EQ403     1              EQ668     1           
    X1028     EQ670     -1             EQ671     -1          
    X1028     EQ677     1                                    
    X1029     EQ201     1              EQ231     -1          
    X1029     EQ377     -1             EQ403     1           
    X1029     EQ538     1              EQ672     -1          
    X1029     EQ673     1                                    
    X1030     EQ201     1              EQ211     -1          
    X1030     EQ231     -1             EQ403     1           
    X1030     EQ538     1              EQ672     1           
    X1030     EQ674     -1                                   
    X1031     EQ403     -1             EQ474     1           
    X1031     EQ475     -1             EQ669     -1          
    X1031     EQ674     1                                    
    X1032     EQ403     1              EQ407     -1          
    X1032     EQ538     1              EQ675     -1          
    X1032     EQ676     1                                    
    X1033     EQ201     1              EQ212     -1          
    X1033     EQ231     -1             EQ403     1           
    X1033     EQ538     1              EQ679     -1          
    X1033     EQ680     1                                    
    X1034     EQ122     1              EQ403     1           
    X1034     EQ792     -1             EQ842     -1          
    X1035     EQ201     1              EQ231     -1          
    X1035     EQ677     1              EQ681     -1          
    X1036     EQ33      -1             EQ208     1           
    X1036     EQ378     -1             EQ461     1           
    X1037     EQ204     2              EQ215     -2          
    X1037     EQ328     1              EQ403     1           
    X1037     EQ683     -1                                   
    X1038     EQ407     1              EQ423     -1          
    X1038     EQ683     1                                    
    X1039     EQ267     4              EQ278     1           
    X1039     EQ403     -4             EQ683     -1          
    X1040     EQ267     4              EQ277     1           
    X1040     EQ403     -4             EQ682     -1          
    X1041     EQ209     1              EQ243     -1          
    X1041     EQ267     1              EQ403     -2          
    X1041     EQ407     -1             EQ478     1           
    X1042     EQ547     1              EQ569     -1          
    X1042     EQ681     1              EQ684     -1          
    X1043     EQ125     1              EQ407     -1          
    X1043     EQ686     -1                                   
    X1044     EQ267     1              EQ403     -2          
    X1044     EQ407     -1             EQ478     2           
    X1044     EQ687     -1                                   
    X1045     EQ86      1              EQ208     -1          
    X1045     EQ378     1              EQ690     -1          
    X1046     EQ217     1              EQ231     -1          
    X1046     EQ547     1              EQ663     -1          
    X1046     EQ690     -1             EQ691     1           
    X1047     EQ201     1              EQ231     -1          
    X1047     EQ403     1              EQ407     -1          
    X1047     EQ538     1              EQ690     1           
    X1047     EQ907     -1                                   
    X1048     EQ784     -1             EQ907     1           
    X1049     EQ403     1              EQ693     1           
    X1049     EQ842     -1             EQ909     -1          
    X1050     EQ785     -1             EQ908     1           
    X1051     EQ786     -1             EQ909     1           
    X1052     EQ403     2              EQ407     -1          
    X1052     EQ472     -1             EQ473     1           
    X1052     EQ692     -1             EQ693     1           
    X1053     EQ121     -1             EQ122     1           
    X1053     EQ403     2              EQ407     -1          
    X1053     EQ472     -1             EQ473     1           
    X1054     EQ48      -1             EQ49      1           
    X1054     EQ403     2              EQ407     -1          
    X1054     EQ472     -1             EQ473     1           
    X1055     EQ403     1              EQ692     1           
    X1055     EQ842     -1             EQ908     -1          
    X1056     EQ49      1              EQ344     1           
    X1056     EQ403     -1             EQ407     1           
    X1056     EQ472     1              EQ473     -1          
    X1056     EQ482     -1             EQ693     -1          
    X1057     EQ403     1              EQ407     -1          
    X1057     EQ472     -1             EQ473     1           
    X1057     EQ685     1              EQ694     -1          
    X1058     EQ241     1              EQ403     -1          
    X1058     EQ407     1              EQ472     1           
    X1058     EQ473     -1             EQ482     -1          
    X1058     EQ642     -1                                   
    X1059     EQ407     -1             EQ592     1           
    X1059     EQ694     1              EQ696     -1          
    X1060     EQ201     1              EQ231     -1          
    X1060     EQ403     1              EQ407     -1          
    X1060     EQ518     1              EQ538     1           
    X1060     EQ876     -1                                   
RHS                                                          
BOUNDS                                                       
 UP BND1      X1        1000                                 
 UP BND1      X2        1000                                 
 UP BND1      X3        1000                                 
 UP BND1      X4        1000                                 
 UP BND1      X5        1000                                 
 UP BND1      X6        1000                                 
 UP BND1      X7        1000                                 
 UP BND1      X8        1000                                 
 UP BND1      X9        1000                                 
 UP BND1      X10       1000                                 
 UP BND1      X11       1000                                 
 UP BND1      X12       1000                                 
 UP BND1      X13       1000                                 
 UP BND1      X14       1000                                 
 UP BND1      X15       1000                                 
 UP BND1      X16       1000                                 
 UP BND1      X17       1000                                 
 UP BND1      X18       1000                                 
 UP BND1      X19       1000                                 
 UP BND1      X20       1000                                 
 UP BND1      X21       1000                                 
 UP BND1      X22       1000                                 
 UP BND1      X23       1000                                 
 UP BND1      X24       1000                                 
 UP BND1      X25       1000                                 
 UP BND1      X26       0                                    
 UP BND1      X27       0                                    
 UP BND1      X28       0                                    
 UP BND1      X29       1000                                 
 UP BND1      X30       1000                                 
 UP BND1      X31       1000                                 
 UP BND1      X32       1000                                 
 UP BND1      X33       1000                                 
 LO BND1      X34       -100                                 
 UP BND1      X34       1000                                 
 LO BND1      X35       -100                                 
 UP BND1      X35       1000                                 
 UP BND1      X36       1000                                 
 UP BND1      X37       1000                                 
 UP BND1      X38       0                                    
 UP BND1      X39       1000                                 
 UP BND1      X40       1000                                 
 UP BND1      X41       1000                                 
 LO BND1      X42       -100                                 
 UP BND1      X42       1000                                 
 UP BND1      X43       1000                                 
 UP BND1      X44       1000                                 
 UP BND1      X45       1000                                 
 UP BND1      X46       1000                                 
 LO BND1      X47       -6.3                                 
 UP BND1      X47       1000                                 
 UP BND1      X48       1000                                 
 UP BND1      X49       0                                    
 UP BND1      X50       0                                    
 UP BND1      X51       1000                                 
 UP BND1      X52       1000                                 
 UP BND1      X53       1000                                 
 UP BND1      X54       1000                                 
 LO BND1      X55       -100                                 
 UP BND1      X55       1000                                 
 LO BND1      X56       -100                                 
 UP BND1      X56       1000                                 
 LO BND1      X57       -100                                 
 UP BND1      X57       1000                                 
 LO BND1      X58       -100                                 
 UP BND1      X58       1000                                 
 UP BND1      X59       1000                                 
 UP BND1      X60       0                                    
 UP BND1      X61       1000                                 
 UP BND1      X62       1000                                 
 UP BND1      X63       0                                    
 UP BND1      X64       1000                                 
 UP BND1      X65       1000                                 
 UP BND1      X66       1000                                 
 UP BND1      X67       0                                    
 UP BND1      X68       0                                    
 UP BND1      X69       1000                                 
 UP BND1      X70       1000                                 
 UP BND1      X71       0                                    
 UP BND1      X72       1000                                 
 LO BND1      X73       -100                                 
 UP BND1      X73       1000                                 
 UP BND1      X74       1000                                 
 LO BND1      X75       -100                                 
 UP BND1      X75       1000                                 
 UP BND1      X76       1000                                 
 UP BND1      X77       1000                                 
 LO BND1      X78       -20                                  
 UP BND1      X78       0                                    
 UP BND1      X79       0                                    
 UP BND1      X80       1000                                 
 UP BND1      X81       1000                                 
 UP BND1      X82       1000                                 
 UP BND1      X83       1000                                 
 UP BND1      X84       0                                    
 LO BND1      X85       -100                                 
 UP BND1      X85       1000                                 
 UP BND1      X86       1000                                 
 UP BND1      X87       0                                    
 UP BND1      X88       1000                                 
 UP BND1      X89       1000                                 
 UP BND1      X90       1000                                 
 UP BND1      X91       0                                    
 LO BND1      X92       -100                                 
 UP BND1      X92       1000                                 
 UP BND1      X93       1000                                 
 UP BND1      X94       1000                                 
 UP BND1      X95       1000                                 
 UP BND1      X96       0                                    
 UP BND1      X97       0                                    
 UP BND1      X98       1000                                 
 UP BND1      X99       1000                                 
 UP BND1      X100      0                                    
 UP BND1      X101      0                                    
 UP BND1      X102      1000                                 
 UP BND1      X103      1000                                 
 LO BND1      X104      -1000                                
 UP BND1      X104      1000                                 
 LO BND1      X105      -1000                                
 UP BND1      X105      1000                                 
 LO BND1      X106      -1000                                
 UP BND1      X106      1000                                 
 LO BND1      X107      -1000                                
 UP BND1      X107      1000                                 
 LO BND1      X108      -1000                                
 UP BND1      X108      1000                                 
 LO BND1      X109      -1000                                
 UP BND1      X109      1000                                 
 LO BND1      X110      -1000                                
 UP BND1      X110      1000                                 
 UP BND1      X111      1000                                 
 UP BND1      X112      1000                                 
 UP BND1      X113      1000                                 
 LO BND1      X114      -1000                                
 UP BND1      X114      1000                                 
 LO BND1      X115      -1000                                
 UP BND1      X115      1000                                 
 LO BND1      X116      -1000                                
 UP BND1      X116      1000                                 
 LO BND1      X117      -1000                                
 UP BND1      X117      1000                                 
 LO BND1      X118      -1000                                
 UP BND1      X118      1000                                 
 UP BND1      X119      1000                                 
 LO BND1      X120      -1000                                
 UP BND1      X120      1000                                 
 LO BND1      X121      -1000                                
 UP BND1      X121      1000                                 
 LO BND1      X122      -1000                                
 UP BND1      X122      1000                                 
 LO BND1      X123      -1000                                
 UP BND1      X123      1000                                 
 LO BND1      X124      -1000                                
 UP BND1      X124      1000                                 
 LO BND1      X125      -1000                                
 UP BND1      X125      1000                                 
 LO BND1      X126      -1000                                
 UP BND1      X126      1000                                 
 UP BND1      X127      1000                                 
 UP BND1      X128      1000                                 
 UP BND1      X129      1000                                 
 UP BND1      X130      1000                                 
 UP BND1      X131      1000                                 
 UP BND1      X132      1000                                 
 UP BND1      X133      1000                                 
 UP BND1      X134      1000                                 
 UP BND1      X135      1000                                 
 UP BND1      X136      1000                                 
 UP BND1      X137      1000                                 
 UP BND1      X138      1000                                 
 UP BND1      X139      1000                                 
 UP BND1      X140      1000                                 
 UP BND1      X141      1000                                 
 UP BND1      X142      1000                                 
 UP BND1      X143      1000                                 
 LO BND1      X144      -1000                                
 UP BND1      X144      1000                                 
 LO BND1      X145      -1000                                
 UP BND1      X145      1000                                 
 UP BND1      X146      1000                                 
 LO BND1      X147      -1000                                
 UP BND1      X147      1000                                 
 UP BND1      X148      1000                                 
 LO BND1      X149      -1000                                
 UP BND1      X149      1000                                 
 UP BND1      X150      1000                                 
 LO BND1      X151      -1000                                
 UP BND1      X151      1000                                 
 UP BND1      X152      1000                                 
 LO BND1      X153      -1000                                
 UP BND1      X153      1000                                 
 LO BND1      X154      -1000                                
 UP BND1      X154      1000                                 
 LO BND1      X155      -1000                                
 UP BND1      X155      1000                                 
 UP BND1      X156      1000                                 
 UP BND1      X157      1000                                 
 UP BND1      X158      1000                                 
 UP BND1      X159      1000                                 
 UP BND1      X160      1000                                 
 UP BND1      X161      1000                                 
 UP BND1      X162      1000                                 
 UP BND1      X163      1000                                 
 UP BND1      X164      1000                                 
 UP BND1      X165      1000                                 
 UP BND1      X166      1000                                 
 UP BND1      X167      1000                                 
 LO BND1      X168      -1000                                
 UP BND1      X168      1000                                 
 LO BND1      X169      -1000                                
 UP BND1      X169      1000                                 
 UP BND1      X170      1000                                 
 UP BND1      X171      1000                                 
 UP BND1      X172      1000                                 
 LO BND1      X173      -1000                                
 UP BND1      X173      1000                                 
 UP BND1      X174      1000                                 
 UP BND1      X175      1000                                 
 UP BND1      X176      1000                                 
 UP BND1      X177      1000                                 
 UP BND1      X178      1000                                 
 UP BND1      X179      1000                                 
 LO BND1      X180      -1000                                
 UP BND1      X180      1000                                 
 LO BND1      X181      -1000                                
 UP BND1      X181      1000                                 
 UP BND1      X182      1000                                 
 LO BND1      X183      -1000                                
 UP BND1      X183      1000                                 
 UP BND1      X184      1000                                 
 LO BND1      X185      -1000                                
 UP BND1      X185      1000                                 
 LO BND1      X186      -1000                                
 UP BND1      X186      1000                                 
 UP BND1      X187      1000                                 
 LO BND1      X188      -1000                                
 UP BND1      X188      1000                                 
 LO BND1      X189      -1000                                
 UP BND1      X189      1000                                 
 LO BND1      X190      -1000                                
 UP BND1      X190      1000                                 
 UP BND1      X191      1000                                 
 LO BND1      X192      -1000                                
 UP BND1      X192      1000                                 
 UP BND1      X193      1000                                 
 UP BND1      X194      1000                                 
 UP BND1      X195      1000                                 
 UP BND1      X196      1000                                 
 UP BND1      X197      1000                                 
 LO BND1      X198      -1000                                
 UP BND1      X198      1000                                 
 LO BND1      X199      -1000                                
 UP BND1      X199      1000                                 
 LO BND1      X200      -1000                                
 UP BND1      X200      1000                                 
 LO BND1      X201      -1000                                
 UP BND1      X201      1000                                 
 LO BND1      X202      -1000                                
 UP BND1      X202      1000                                 
 LO BND1      X203      -1000                                
 UP BND1      X203      1000                                 
 LO BND1      X204      -1000                                
 UP BND1      X204      1000                                 
 UP BND1      X205      1000                                 
 UP BND1      X206      1000                                 
 LO BND1      X207      -1000                                
 UP BND1      X207      1000                                 
 UP BND1      X208      1000                                 
 LO BND1      X209      -1000                                
 UP BND1      X209      1000                                 
 LO BND1      X210      -1000                                
 UP BND1      X210      1000                                 
 LO BND1      X211      -1000                                
 UP BND1      X211      1000                                 
 LO BND1      X212      -1000                                
 UP BND1      X212      1000                                 
 LO BND1      X213      -1000                                
 UP BND1      X213      1000                                 
 UP BND1      X214      1000                                 
 LO BND1      X215      -1000                                
 UP BND1      X215      1000                                 
 UP BND1      X216      1000                                 
 LO BND1      X217      -1000                                
 UP BND1      X217      1000                                 
 UP BND1      X218      1000                                 
 UP BND1      X219      1000                                 
 UP BND1      X220      1000                                 
 UP BND1      X221      1000                                 
 UP BND1      X222      1000                                 
 LO BND1      X223      -1000                                
 UP BND1      X223      1000                                 
 UP BND1      X224      1000                                 
 UP BND1      X225      1000                                 
 UP BND1      X226      1000                                 
 UP BND1      X227      1000                                 
 LO BND1      X228      -1000                                
 UP BND1      X228      1000                                 
 LO BND1      X229      -1000                                
 UP BND1      X229      1000                                 
 LO BND1      X230      -1000                                
 UP BND1      X230      1000                                 
 UP BND1      X231      1000                                 
 UP BND1      X232      1000                                 
 UP BND1      X233      1000                                 
 UP BND1      X234      1000                                 
 UP BND1      X235      1000                                 
 UP BND1      X236      1000                                 
 UP BND1      X237      1000                                 
 LO BND1      X238      -1000                                
 UP BND1      X238      1000                                 
 LO BND1      X239      -1000                                
 UP BND1      X239      1000                                 
 LO BND1      X240      -1000                                
 UP BND1      X240      1000                                 
 UP BND1      X241      1000                                 
 LO BND1      X242      -1000                                
 UP BND1      X242      1000                                 
 LO BND1      X243      -1000                                
 UP BND1      X243      1000                                 
 LO BND1      X244      -1000                                
 UP BND1      X244      1000                                 
 LO BND1      X245      -1000                                
 UP BND1      X245      1000                                 
 LO BND1      X246      -1000                                
 UP BND1      X246      1000                                 
 LO BND1      X247      -1000                                
 UP BND1      X247      1000                                 
 LO BND1      X248      -1000                                
 UP BND1      X248      1000                                 
 UP BND1      X249      1000                                 
 LO BND1      X250      -1000                                
 UP BND1      X250      1000                                 
 LO BND1      X251      -1000                                
 UP BND1      X251      1000                                 
 UP BND1      X252      1000                                 
 UP BND1      X253      1000                                 
 UP BND1      X254      1000                                 
 LO BND1      X255      -1000                                
 UP BND1      X255      1000                                 
 UP BND1      X256      1000                                 
 UP BND1      X257      1000                                 
 UP BND1      X258      1000                                 
 UP BND1      X259      1000                                 
 LO BND1      X260      -1000                                
 UP BND1      X260      1000                                 
 UP BND1      X261      1000                                 
 UP BND1      X262      1000                                 
 UP BND1      X263      1000                                 
 UP BND1      X264      1000                                 
 LO BND1      X265      -1000                                
 UP BND1      X265      1000                                 
 UP BND1      X266      1000                                 
 UP BND1      X267      1000                                 
 LO BND1      X268      -1000                                
 UP BND1      X268      1000                                 
 LO BND1      X269      -1000                                
 UP BND1      X269      1000                                 
 UP BND1      X270      1000                                 
 UP BND1      X271      1000                                 
 UP BND1      X272      1000                                 
 LO BND1      X273      -1000                                
 UP BND1      X273      1000                                 
 LO BND1      X274      -1000                                
 UP BND1      X274      1000                                 
 UP BND1      X275      1000                                 
 UP BND1      X276      1000                                 
 UP BND1      X277      1000                                 
 UP BND1      X278      1000                                 
 UP BND1      X279      1000                                 
 UP BND1      X280      1000                                 
 UP BND1      X281      1000                                 
 UP BND1      X282      1000                                 
 LO BND1      X283      -1000                                
 UP BND1      X283      1000                                 
 UP BND1      X284      1000                                 
 UP BND1      X285      1000                                 
 UP BND1      X286      1000                                 
 LO BND1      X287      -1000                                
 UP BND1      X287      1000                                 
 UP BND1      X288      1000                                 
 LO BND1      X289      -1000                                
 UP BND1      X289      1000                                 
 LO BND1      X290      -1000                                
 UP BND1      X290      1000                                 
 LO BND1      X291      -1000                                
 UP BND1      X291      1000                                 
 UP BND1      X292      1000                                 
 UP BND1      X293      1000                                 
 UP BND1      X294      1000                                 
 UP BND1      X295      1000                                 
 UP BND1      X296      1000                                 
 UP BND1      X297      1000                                 
 LO BND1      X298      -1000                                
 UP BND1      X298      1000                                 
 LO BND1      X299      -1000                                
 UP BND1      X299      1000                                 
 UP BND1      X300      1000                                 
 UP BND1      X301      1000                                 
 UP BND1      X302      1000                                 
 UP BND1      X303      1000                                 
 UP BND1      X304      1000                                 
 LO BND1      X305      -1000                                
 UP BND1      X305      1000                                 
 LO BND1      X306      -1000                                
 UP BND1      X306      1000                                 
 LO BND1      X307      -1000                                
 UP BND1      X307      1000                                 
 UP BND1      X308      1000                                 
 UP BND1      X309      1000                                 
 UP BND1      X310      1000                                 
 UP BND1      X311      1000                                 
 UP BND1      X312      1000                                 
 UP BND1      X313      1000                                 
 UP BND1      X314      1000                                 
 UP BND1      X315      1000                                 
 LO BND1      X316      -1000                                
 UP BND1      X316      1000                                 
 UP BND1      X317      1000                                 
 LO BND1      X318      -1000                                
 UP BND1      X318      1000                                 
 UP BND1      X319      1000                                 
 UP BND1      X320      1000                                 
 UP BND1      X321      1000                                 
 UP BND1      X322      1000                                 
 LO BND1      X323      -1000                                
 UP BND1      X323      1000                                 
 LO BND1      X324      -1000                                
 UP BND1      X324      1000                                 
 LO BND1      X325      -1000                                
 UP BND1      X325      1000                                 
 LO BND1      X326      -1000                                
 UP BND1      X326      1000                                 
 LO BND1      X327      -1000                                
 UP BND1      X327      1000                                 
 LO BND1      X328      -1000                                
 UP BND1      X328      1000                                 
 LO BND1      X329      -1000                                
 UP BND1      X329      1000                                 
 LO BND1      X330      -1000                                
 UP BND1      X330      1000                                 
 LO BND1      X331      -1000                                
 UP BND1      X331      1000                                 
 UP BND1      X332      1000                                 
 UP BND1      X333      1000                                 
 LO BND1      X334      -1000                                
 UP BND1      X334      1000                                 
 UP BND1      X335      1000                                 
 UP BND1      X336      1000                                 
 UP BND1      X337      1000                                 
 LO BND1      X338      -1000                                
 UP BND1      X338      1000                                 
 LO BND1      X339      -1000                                
 UP BND1      X339      1000                                 
 UP BND1      X340      1000                                 
 UP BND1      X341      1000                                 
 UP BND1      X342      1000                                 
 LO BND1      X343      -1000                                
 UP BND1      X343      1000                                 
 UP BND1      X344      1000                                 
 UP BND1      X345      1000                                 
 UP BND1      X346      1000                                 
 UP BND1      X347      1000                                 
 UP BND1      X348      1000                                 
 UP BND1      X349      1000                                 
 UP BND1      X350      1000                                 
 UP BND1      X351      1000                                 
 UP BND1      X352      1000                                 
 UP BND1      X353      1000                                 
 UP BND1      X354      1000                                 
 UP BND1      X355      1000                                 
 UP BND1      X356      1000                                 
 LO BND1      X357      -1000                                
 UP BND1      X357      1000                                 
 UP BND1      X358      1000                                 
 UP BND1      X359      1000                                 
 LO BND1      X360      -1000                                
 UP BND1      X360      1000                                 
 LO BND1      X361      -1000                                
 UP BND1      X361      1000                                 
 LO BND1      X362      -1000                                
 UP BND1      X362      1000                                 
 LO BND1      X363      -1000                                
 UP BND1      X363      1000                                 
 UP BND1      X364      1000                                 
 LO BND1      X365      -1000                                
 UP BND1      X365      1000                                 
 UP BND1      X366      1000                                 
 LO BND1      X367      -1000                                
 UP BND1      X367      1000                                 
 UP BND1      X368      1000                                 
 UP BND1      X369      1000                                 
 LO BND1      X370      -1000                                
 UP BND1      X370      1000                                 
 UP BND1      X371      1000                                 
 UP BND1      X372      1000                                 
 UP BND1      X373      1000                                 
 UP BND1      X374      1000                                 
 UP BND1      X375      1000                                 
 UP BND1      X376      1000                                 
 LO BND1      X377      -1000                                
 UP BND1      X377      1000                                 
 UP BND1      X378      1000                                 
 UP BND1      X379      1000                                 
 LO BND1      X380      -1000                                
 UP BND1      X380      1000                                 
 UP BND1      X381      1000                                 
 LO BND1      X382      -1000                                
 UP BND1      X382      1000                                 
 UP BND1      X383      1000                                 
 UP BND1      X384      1000                                 
 UP BND1      X385      1000                                 
 UP BND1      X386      1000                                 
 UP BND1      X387      1000                                 
 UP BND1      X388      1000                                 
 LO BND1      X389      -1000                                
 UP BND1      X389      1000                                 
 UP BND1      X390      1000                                 
 UP BND1      X391      1000                                 
 LO BND1      X392      -1000                                
 UP BND1      X392      1000                                 
 LO BND1      X393      -1000                                
 UP BND1      X393      1000                                 
 UP BND1      X394      1000                                 
 UP BND1      X395      1000                                 
 UP BND1      X396      1000                                 
 UP BND1      X397      1000                                 
 UP BND1      X398      1000                                 
 LO BND1      X399      -1000                                
 UP BND1      X399      1000                                 
 UP BND1      X400      1000                                 
 UP BND1      X401      1000                                 
 LO BND1      X402      -1000                                
 UP BND1      X402      1000                                 
 LO BND1      X403      -1000                                
 UP BND1      X403      1000                                 
 UP BND1      X404      1000                                 
 LO BND1      X405      -1000                                
 UP BND1      X405      1000                                 
 LO BND1      X406      -1000                                
 UP BND1      X406      1000                                 
 UP BND1      X407      1000                                 
 UP BND1      X408      1000                                 
 UP BND1      X409      1000                                 
 UP BND1      X410      1000                                 
 UP BND1      X411      1000                                 
 UP BND1      X412      1000                                 
 UP BND1      X413      1000                                 
 UP BND1      X414      1000                                 
 UP BND1      X415      1000                                 
 UP BND1      X416      1000                                 
 UP BND1      X417      1000                                 
 LO BND1      X418      -1000                                
 UP BND1      X418      1000                                 
 LO BND1      X419      -1000                                
 UP BND1      X419      1000                                 
 UP BND1      X420      1000                                 
 UP BND1      X421      1000                                 
 UP BND1      X422      1000                                 
 LO BND1      X423      -1000                                
 UP BND1      X423      1000                                 
 UP BND1      X424      1000                                 
 UP BND1      X425      1000                                 
 UP BND1      X426      1000                                 
 UP BND1      X427      1000                                 
 UP BND1      X428      1000                                 
 UP BND1      X429      1000                                 
 UP BND1      X430      1000                                 
 UP BND1      X431      1000                                 
 UP BND1      X432      1000                                 
 UP BND1      X433      1000                                 
 UP BND1      X434      1000                                 
 UP BND1      X435      1000                                 
 LO BND1      X436      -1000                                
 UP BND1      X436      1000                                 
 UP BND1      X437      1000                                 
 LO BND1      X438      -1000                                
 UP BND1      X438      1000                                 
 LO BND1      X439      -1000                                
 UP BND1      X439      1000                                 
 LO BND1      X440      -1000                                
 UP BND1      X440      1000                                 
 LO BND1      X441      -1000                                
 UP BND1      X441      1000                                 
 LO BND1      X442      -1000                                
 UP BND1      X442      1000                                 
 LO BND1      X443      -1000                                
 UP BND1      X443      1000                                 
 UP BND1      X444      1000                                 
 UP BND1      X445      1000                                 
 UP BND1      X446      1000                                 
 LO BND1      X447      -1000                                
 UP BND1      X447      1000                                 
 UP BND1      X448      1000                                 
 LO BND1      X449      -1000                                
 UP BND1      X449      1000                                 
 UP BND1      X450      1000                                 
 UP BND1      X451      1000                                 
 UP BND1      X452      1000                                 
 UP BND1      X453      1000                                 
 UP BND1      X454      1000                                 
 UP BND1      X455      1000                                 
 UP BND1      X456      1000                                 
 LO BND1      X457      -1000                                
 UP BND1      X457      1000                                 
 UP BND1      X458      1000                                 
 UP BND1      X459      1000                                 
 UP BND1      X460      1000                                 
 UP BND1      X461      1000                                 
 UP BND1      X462      1000                                 
 UP BND1      X463      1000                                 
 UP BND1      X464      1000                                 
 UP BND1      X465      1000                                 
 UP BND1      X466      1000                                 
 UP BND1      X467      1000                                 
 LO BND1      X468      -1000                                
 UP BND1      X468      1000                                 
 UP BND1      X469      1000                                 
 LO BND1      X470      -1000                                
 UP BND1      X470      1000                                 
 LO BND1      X471      -1000                                
 UP BND1      X471      1000                                 
 UP BND1      X472      1000                                 
 UP BND1      X473      1000                                 
 UP BND1      X474      1000                                 
 UP BND1      X475      1000                                 
 UP BND1      X476      1000                                 
 LO BND1      X477      -1000                                
 UP BND1      X477      1000                                 
 UP BND1      X478      1000                                 
 UP BND1      X479      1000                                 
 LO BND1      X480      -1000                                
 UP BND1      X480      1000                                 
 UP BND1      X481      1000                                 
 UP BND1      X482      1000                                 
 LO BND1      X483      -1000                                
 UP BND1      X483      1000                                 
 UP BND1      X484      1000                                 
 UP BND1      X485      1000                                 
 UP BND1      X486      1000                                 
 UP BND1      X487      1000                                 
 UP BND1      X488      1000                                 
 UP BND1      X489      1000                                 
 UP BND1      X490      1000                                 
 UP BND1      X491      1000                                 
 LO BND1      X492      -1000                                
 UP BND1      X492      1000                                 
 LO BND1      X493      -1000                                
 UP BND1      X493      1000                                 
 UP BND1      X494      1000                                 
 UP BND1      X495      1000                                 
 UP BND1      X496      1000                                 
 LO BND1      X497      -1000                                
 UP BND1      X497      1000                                 
 UP BND1      X498      1000                                 
 LO BND1      X499      -1000                                
 UP BND1      X499      1000                                 
 LO BND1      X500      -1000                                
 UP BND1      X500      1000                                 
 LO BND1      X501      -1000                                
 UP BND1      X501      1000                                 
 LO BND1      X502      -1000                                
 UP BND1      X502      1000                                 
 UP BND1      X503      1000                                 
 LO BND1      X504      -1000                                
 UP BND1      X504      1000                                 
 LO BND1      X505      -1000                                
 UP BND1      X505      1000                                 
 UP BND1      X506      1000                                 
 UP BND1      X507      1000                                 
 LO BND1      X508      -1000                                
 UP BND1      X508      1000                                 
 UP BND1      X509      1000                                 
 LO BND1      X510      -1000                                
 UP BND1      X510      1000                                 
 LO BND1      X511      -1000                                
 UP BND1      X511      1000                                 
 UP BND1      X512      1000                                 
 LO BND1      X513      -1000                                
 UP BND1      X513      1000                                 
 LO BND1      X514      -1000                                
 UP BND1      X514      1000                                 
 UP BND1      X515      1000                                 
 UP BND1      X516      1000                                 
 UP BND1      X517      1000                                 
 LO BND1      X518      -1000                                
 UP BND1      X518      1000                                 
 UP BND1      X519      1000                                 
 UP BND1      X520      1000                                 
 UP BND1      X521      1000                                 
 UP BND1      X522      1000                                 
 UP BND1      X523      1000                                 
 UP BND1      X524      1000                                 
 LO BND1      X525      -1000                                
 UP BND1      X525      1000                                 
 UP BND1      X526      1000                                 
 UP BND1      X527      1000                                 
 LO BND1      X528      -1000                                
 UP BND1      X528      1000                                 
 UP BND1      X529      1000                                 
 UP BND1      X530      1000                                 
 UP BND1      X531      1000                                 
 UP BND1      X532      1000                                 
 UP BND1      X533      1000                                 
 UP BND1      X534      1000                                 
 LO BND1      X535      -1000                                
 UP BND1      X535      1000                                 
 UP BND1      X536      1000                                 
 LO BND1      X537      -1000                                
 UP BND1      X537      1000                                 
 LO BND1      X538      -1000                                
 UP BND1      X538      1000                                 
 LO BND1      X539      -1000                                
 UP BND1      X539      1000                                 
 UP BND1      X540      1000                                 
 LO BND1      X541      -1000                                
 UP BND1      X541      1000                                 
 UP BND1      X542      1000                                 
 UP BND1      X543      1000                                 
 LO BND1      X544      -1000                                
 UP BND1      X544      1000                                 
 LO BND1      X545      -1000                                
 UP BND1      X545      1000                                 
 UP BND1      X546      1000                                 
 UP BND1      X547      1000                                 
 LO BND1      X548      -1000                                
 UP BND1      X548      1000                                 
 LO BND1      X549      -1000                                
 UP BND1      X549      1000                                 
 LO BND1      X550      -1000                                
 UP BND1      X550      1000                                 
 UP BND1      X551      1000                                 
 UP BND1      X552      1000                                 
 LO BND1      X553      -1000                                
 UP BND1      X553      1000                                 
 LO BND1      X554      -1000                                
 UP BND1      X554      1000                                 
 UP BND1      X555      1000                                 
 UP BND1      X556      1000                                 
 UP BND1      X557      1000                                 
 UP BND1      X558      1000                                 
 UP BND1      X559      1000                                 
 UP BND1      X560      1000                                 
 LO BND1      X561      -1000                                
 UP BND1      X561      1000                                 
 UP BND1      X562      1000                                 
 UP BND1      X563      1000                                 
 UP BND1      X564      1000                                 
 UP BND1      X565      1000                                 
 UP BND1      X566      1000                                 
 UP BND1      X567      1000                                 
 UP BND1      X568      1000                                 
 LO BND1      X569      -1000                                
 UP BND1      X569      1000                                 
 LO BND1      X570      -1000                                
 UP BND1      X570      1000                                 
 LO BND1      X571      -1000                                
 UP BND1      X571      1000                                 
 LO BND1      X572      -1000                                
 UP BND1      X572      1000                                 
 LO BND1      X573      -1000                                
 UP BND1      X573      1000                                 
 LO BND1      X574      -1000                                
 UP BND1      X574      1000                                 
 UP BND1      X575      1000                                 
 UP BND1      X576      1000                                 
 UP BND1      X577      1000                                 
 UP BND1      X578      1000                                 
 UP BND1      X579      1000                                 
 UP BND1      X580      1000                                 
 UP BND1      X581      1000                                 
 LO BND1      X582      -1000                                
 UP BND1      X582      1000                                 
 UP BND1      X583      1000                                 
 LO BND1      X584      -1000                                
 UP BND1      X584      1000                                 
 LO BND1      X585      -1000                                
 UP BND1      X585      1000                                 
 LO BND1      X586      -1000                                
 UP BND1      X586      1000                                 
 UP BND1      X587      1000                                 
 UP BND1      X588      1000                                 
 UP BND1      X589      1000                                 
 UP BND1      X590      1000                                 
 UP BND1      X591      1000                                 
 UP BND1      X592      1000                                 
 UP BND1      X593      1000                                 
 UP BND1      X594      1000                                 
 UP BND1      X595      1000                                 
 LO BND1      X596      -1000                                
 UP BND1      X596      1000                                 
 UP BND1      X597      1000                                 
 UP BND1      X598      1000                                 
 UP BND1      X599      1000                                 
 UP BND1      X600      1000                                 
 UP BND1      X601      1000                                 
 UP BND1      X602      1000                                 
 UP BND1      X603      1000                                 
 UP BND1      X604      1000                                 
 UP BND1      X605      1000                                 
 UP BND1      X606      1000                                 
 UP BND1      X607      1000                                 
 UP BND1      X608      1000                                 
 LO BND1      X609      -1000                                
 UP BND1      X609      1000                                 
 LO BND1      X610      -1000                                
 UP BND1      X610      1000                                 
 LO BND1      X611      -1000                                
 UP BND1      X611      1000                                 
 UP BND1      X612      1000                                 
 LO BND1      X613      -1000                                
 UP BND1      X613      1000                                 
 UP BND1      X614      1000                                 
 LO BND1      X615      -1000                                
 UP BND1      X615      1000                                 
 UP BND1      X616      1000                                 
 LO BND1      X617      -1000                                
 UP BND1      X617      1000                                 
 LO BND1      X618      -1000                                
 UP BND1      X618      1000                                 
 LO BND1      X619      -1000                                
 UP BND1      X619      1000                                 
 LO BND1      X620      -1000                                
 UP BND1      X620      1000                                 
 UP BND1      X621      1000                                 
 UP BND1      X622      1000                                 
 UP BND1      X623      1000                                 
 UP BND1      X624      1000                                 
 LO BND1      X625      -1000                                
 UP BND1      X625      1000                                 
 UP BND1      X626      1000                                 
 UP BND1      X627      1000                                 
 LO BND1      X628      -1000                                
 UP BND1      X628      1000                                 
 LO BND1      X629      -1000                                
 UP BND1      X629      1000                                 
 UP BND1      X630      1000                                 
 UP BND1      X631      1000                                 
 LO BND1      X632      -1000                                
 UP BND1      X632      1000                                 
 UP BND1      X633      1000                                 
 LO BND1      X634      -1000                                
 UP BND1      X634      1000                                 
 LO BND1      X635      -1000                                
 UP BND1      X635      1000                                 
 LO BND1      X636      -1000                                
 UP BND1      X636      1000                                 
 UP BND1      X637      1000                                 
 UP BND1      X638      1000                                 
 LO BND1      X639      -1000                                
 UP BND1      X639      1000                                 
 UP BND1      X640      1000                                 
 LO BND1      X641      -1000                                
 UP BND1      X641      1000                                 
 LO BND1      X642      -1000                                
 UP BND1      X642      1000                                 
 UP BND1      X643      1000                                 
 UP BND1      X644      1000                                 
 UP BND1      X645      1000                                 
 UP BND1      X646      1000                                 
 UP BND1      X647      1000                                 
 UP BND1      X648      1000                                 
 UP BND1      X649      1000                                 
 UP BND1      X650      1000                                 
 UP BND1      X651      1000                                 
 UP BND1      X652      1000                                 
 LO BND1      X653      -1000                                
 UP BND1      X653      1000                                 
 LO BND1      X654      -1000                                
 UP BND1      X654      1000                                 
 LO BND1      X655      -1000                                
 UP BND1      X655      1000                                 
 LO BND1      X656      -1000                                
 UP BND1      X656      1000                                 
 UP BND1      X657      1000                                 
 UP BND1      X658      1000                                 
 LO BND1      X659      -1000                                
 UP BND1      X659      1000                                 
 UP BND1      X660      1000                                 
 UP BND1      X661      1000                                 
 UP BND1      X662      1000                                 
 LO BND1      X663      -1000                                
 UP BND1      X663      1000                                 
 UP BND1      X664      1000                                 
 UP BND1      X665      1000                                 
 LO BND1      X666      -1000                                
 UP BND1      X666      1000                                 
 LO BND1      X667      -1000                                
 UP BND1      X667      1000                                 
 LO BND1      X668      -1000                                
 UP BND1      X668      1000                                 
 UP BND1      X669      1000                                 
 UP BND1      X670      1000                                 
 UP BND1      X671      1000                                 
 LO BND1      X672      -1000                                
 UP BND1      X672      1000                                 
 UP BND1      X673      1000                                 
 UP BND1      X674      1000                                 
 UP BND1      X675      1000                                 
 LO BND1      X676      -1000                                
 UP BND1      X676      1000                                 
 UP BND1      X677      1000                                 
 UP BND1      X678      1000                                 
 UP BND1      X679      1000                                 
 LO BND1      X680      -1000                                
 UP BND1      X680      1000                                 
 UP BND1      X681      1000                                 
 LO BND1      X682      -1000                                
 UP BND1      X682      1000                                 
 LO BND1      X683      -1000                                
 UP BND1      X683      1000                                 
 UP BND1      X684      1000                                 
 UP BND1      X685      1000                                 
 UP BND1      X686      1000                                 
 UP BND1      X687      1000                                 
 UP BND1      X688      1000                                 
 LO BND1      X689      -1000                                
 UP BND1      X689      1000                                 
 UP BND1      X690      1000                                 
 UP BND1      X691      1000                                 
 LO BND1      X692      -1000                                
 UP BND1      X692      1000                                 
 UP BND1      X693      1000                                 
 UP BND1      X694      1000                                 
 LO BND1      X695      -1000                                
 UP BND1      X695      1000                                 
 UP BND1      X696      1000                                 
 LO BND1      X697      -1000                                
 UP BND1      X697      1000                                 
 UP BND1      X698      1000                                 
 UP BND1      X699      1000                                 
 UP BND1      X700      1000                                 
 LO BND1      X701      -1000                                
 UP BND1      X701      1000                                 
 LO BND1      X702      -1000                                
 UP BND1      X702      1000                                 
 LO BND1      X703      -1000                                
 UP BND1      X703      1000                                 
 LO BND1      X704      -1000                                
 UP BND1      X704      1000                                 
 UP BND1      X705      1000                                 
 LO BND1      X706      -1000                                
 UP BND1      X706      1000                                 
 UP BND1      X707      1000                                 
 UP BND1      X708      1000                                 
 LO BND1      X709      -1000                                
 UP BND1      X709      1000                                 
 LO BND1      X710      -1000                                
 UP BND1      X710      1000                                 
 UP BND1      X711      1000                                 
 UP BND1      X712      1000                                 
 LO BND1      X713      -1000                                
 UP BND1      X713      1000                                 
 LO BND1      X714      -1000                                
 UP BND1      X714      1000                                 
 UP BND1      X715      1000                                 
 UP BND1      X716      1000                                 
 UP BND1      X717      1000                                 
 UP BND1      X718      1000                                 
 UP BND1      X719      1000                                 
 UP BND1      X720      1000                                 
 UP BND1      X721      1000                                 
 UP BND1      X722      1000                                 
 LO BND1      X723      -1000                                
 UP BND1      X723      1000                                 
 LO BND1      X724      -1000                                
 UP BND1      X724      1000                                 
 LO BND1      X725      -1000                                
 UP BND1      X725      1000                                 
 LO BND1      X726      -1000                                
 UP BND1      X726      1000                                 
 LO BND1      X727      -1000                                
 UP BND1      X727      1000                                 
 LO BND1      X728      -1000                                
 UP BND1      X728      1000                                 
 LO BND1      X729      -1000                                
 UP BND1      X729      1000                                 
 LO BND1      X730      -1000                                
 UP BND1      X730      1000                                 
 LO BND1      X731      -1000                                
 UP BND1      X731      1000                                 
 UP BND1      X732      1000                                 
 UP BND1      X733      1000                                 
 LO BND1      X734      -1000                                
 UP BND1      X734      1000                                 
 LO BND1      X735      -1000                                
 UP BND1      X735      1000                                 
 LO BND1      X736      -1000                                
 UP BND1      X736      1000                                 
 UP BND1      X737      1000                                 
 LO BND1      X738      -1000                                
 UP BND1      X738      1000                                 
 UP BND1      X739      1000                                 
 UP BND1      X740      1000                                 
 LO BND1      X741      -1000                                
 UP BND1      X741      1000                                 
 LO BND1      X742      -1000                                
 UP BND1      X742      1000                                 
 UP BND1      X743      1000                                 
 UP BND1      X744      1000                                 
 UP BND1      X745      1000                                 
 UP BND1      X746      1000                                 
 UP BND1      X747      1000                                 
 UP BND1      X748      1000                                 
 UP BND1      X749      1000                                 
 UP BND1      X750      1000                                 
 UP BND1      X751      1000                                 
 UP BND1      X752      1000                                 
 UP BND1      X753      1000                                 
 UP BND1      X754      1000                                 
 UP BND1      X755      1000                                 
 UP BND1      X756      1000                                 
 UP BND1      X757      1000                                 
 UP BND1      X758      1000                                 
 UP BND1      X759      1000                                 
 UP BND1      X760      1000                                 
 UP BND1      X761      1000                                 
 UP BND1      X762      1000                                 
 UP BND1      X763      1000                                 
 LO BND1      X764      -1000                                
 UP BND1      X764      1000                                 
 LO BND1      X765      -1000                                
 UP BND1      X765      1000                                 
 UP BND1      X766      1000                                 
 UP BND1      X767      1000                                 
 UP BND1      X768      1000                                 
 UP BND1      X769      1000                                 
 LO BND1      X770      -1000                                
 UP BND1      X770      1000                                 
 UP BND1      X771      1000                                 
 LO BND1      X772      -1000                                
 UP BND1      X772      1000                                 
 LO BND1      X773      -1000                                
 UP BND1      X773      1000                                 
 LO BND1      X774      -1000                                
 UP BND1      X774      1000                                 
 UP BND1      X775      1000                                 
 UP BND1      X776      1000                                 
 UP BND1      X777      1000                                 
 LO BND1      X778      -1000                                
 UP BND1      X778      1000                                 
 UP BND1      X779      1000                                 
 UP BND1      X780      1000                                 
 UP BND1      X781      1000                                 
 UP BND1      X782      1000                                 
 UP BND1      X783      1000                                 
 UP BND1      X784      1000                                 
 UP BND1      X785      1000                                 
 LO BND1      X786      -1000                                
 UP BND1      X786      1000                                 
 UP BND1      X787      1000                                 
 LO BND1      X788      -1000                                
 UP BND1      X788      1000                                 
 LO BND1      X789      -1000                                
 UP BND1      X789      1000                                 
 UP BND1      X790      1000                                 
 UP BND1      X791      1000                                 
 LO BND1      X792      -1000                                
 UP BND1      X792      1000                                 
 LO BND1      X793      -1000                                
 UP BND1      X793      1000                                 
 LO BND1      X794      -1000                                
 UP BND1      X794      1000                                 
 UP BND1      X795      1000                                 
 UP BND1      X796      1000                                 
 UP BND1      X797      1000                                 
 LO BND1      X798      -1000                                
 UP BND1      X798      1000                                 
 LO BND1      X799      -1000                                
 UP BND1      X799      1000                                 
 LO BND1      X800      -1000                                
 UP BND1      X800      1000                                 
 UP BND1      X801      1000                                 
 UP BND1      X802      1000                                 
 UP BND1      X803      1000                                 
 UP BND1      X804      1000                                 
 UP BND1      X805      1000                                 
 UP BND1      X806      1000                                 
 UP BND1      X807      1000                                 
 UP BND1      X808      1000                                 
 UP BND1      X809      1000                                 
 UP BND1      X810      1000                                 
 UP BND1      X811      1000                                 
 UP BND1      X812      1000                                 
 LO BND1      X813      -1000                                
 UP BND1      X813      1000                                 
 LO BND1      X814      -1000                                
 UP BND1      X814      1000                                 
 UP BND1      X815      1000                                 
 UP BND1      X816      1000                                 
 UP BND1      X817      1000                                 
 UP BND1      X818      1000                                 
 UP BND1      X819      1000                                 
 UP BND1      X820      1000                                 
 UP BND1      X821      1000                                 
 UP BND1      X822      1000                                 
 LO BND1      X823      -1000                                
 UP BND1      X823      1000                                 
 LO BND1      X824      -1000                                
 UP BND1      X824      1000                                 
 LO BND1      X825      -1000                                
 UP BND1      X825      1000                                 
 LO BND1      X826      -1000                                
 UP BND1      X826      1000                                 
 LO BND1      X827      -1000                                
 UP BND1      X827      1000                                 
 LO BND1      X828      -1000                                
 UP BND1      X828      1000                                 
 UP BND1      X829      1000                                 
 LO BND1      X830      -1000                                
 UP BND1      X830      1000                                 
 LO BND1      X831      -1000                                
 UP BND1      X831      1000                                 
 LO BND1      X832      -1000                                
 UP BND1      X832      1000                                 
 UP BND1      X833      1000                                 
 UP BND1      X834      1000                                 
 UP BND1      X835      1000                                 
 UP BND1      X836      1000                                 
 UP BND1      X837      1000                                 
 UP BND1      X838      1000                                 
 UP BND1      X839      1000                                 
 UP BND1      X840      1000                                 
 UP BND1      X841      1000                                 
 UP BND1      X842      1000                                 
 UP BND1      X843      1000                                 
 UP BND1      X844      1000                                 
 UP BND1      X845      1000                                 
 UP BND1      X846      1000                                 
 UP BND1      X847      1000                                 
 UP BND1      X848      1000                                 
 UP BND1      X849      1000                                 
 UP BND1      X850      1000                                 
 UP BND1      X851      1000                                 
 UP BND1      X852      1000                                 
 LO BND1      X853      -1000                                
 UP BND1      X853      1000                                 
 UP BND1      X854      1000                                 
 UP BND1      X855      1000                                 
 UP BND1      X856      1000                                 
 UP BND1      X857      1000                                 
 UP BND1      X858      1000                                 
 UP BND1      X859      1000                                 
 UP BND1      X860      1000                                 
 UP BND1      X861      1000                                 
 LO BND1      X862      -1000                                
 UP BND1      X862      1000                                 
 UP BND1      X863      1000                                 
 UP BND1      X864      1000                                 
 LO BND1      X865      -1000                                
 UP BND1      X865      1000                                 
 LO BND1      X866      -1000                                
 UP BND1      X866      1000                                 
 LO BND1      X867      -1000                                
 UP BND1      X867      1000                                 
 LO BND1      X868      -1000                                
 UP BND1      X868      1000                                 
 UP BND1      X869      1000                                 
 UP BND1      X870      1000                                 
 UP BND1      X871      1000                                 
 UP BND1      X872      1000                                 
 UP BND1      X873      1000                                 
 UP BND1      X874      1000                                 
 LO BND1      X875      -1000                                
 UP BND1      X875      1000                                 
 UP BND1      X876      1000                                 
 UP BND1      X877      1000                                 
 UP BND1      X878      1000                                 
 UP BND1      X879      1000                                 
 LO BND1      X880      -1000                                
 UP BND1      X880      1000                                 
 LO BND1      X881      -1000                                
 UP BND1      X881      1000                                 
 UP BND1      X882      1000                                 
 UP BND1      X883      1000                                 
 UP BND1      X884      1000                                 
 UP BND1      X885      1000                                 
 UP BND1      X886      1000                                 
 UP BND1      X887      1000                                 
 UP BND1      X888      1000                                 
 UP BND1      X889      1000                                 
 LO BND1      X890      -1000                                
 UP BND1      X890      1000                                 
 UP BND1      X891      1000                                 
 UP BND1      X892      1000                                 
 LO BND1      X893      -1000                                
 UP BND1      X893      1000                                 
 UP BND1      X894      1000                                 
 UP BND1      X895      1000                                 
 UP BND1      X896      1000                                 
 UP BND1      X897      1000                                 
 UP BND1      X898      1000                                 
 UP BND1      X899      1000                                 
 UP BND1      X900      1000                                 
 UP BND1      X901      1000                                 
 LO BND1      X902      -1000                                
 UP BND1      X902      1000                                 
 LO BND1      X903      -1000                                
 UP BND1      X903      1000                                 
 LO BND1      X904      -1000                                
 UP BND1      X904      1000                                 
 UP BND1      X905      1000                                 
 UP BND1      X906      1000                                 
 UP BND1      X907      1000                                 
 UP BND1      X908      1000                                 
 UP BND1      X909      1000                                 
 LO BND1      X910      -1000                                
 UP BND1      X910      1000                                 
 UP BND1      X911      1000                                 
 UP BND1      X912      1000                                 
 UP BND1      X913      1000                                 
 UP BND1      X914      1000                                 
 UP BND1      X915      1000                                 
 UP BND1      X916      1000                                 
 LO BND1      X917      -1000                                
 UP BND1      X917      1000                                 
 UP BND1      X918      1000                                 
 LO BND1      X919      -1000                                
 UP BND1      X919      1000                                 
 UP BND1      X920      1000                                 
 LO BND1      X921      -1000                                
 UP BND1      X921      1000                                 
 UP BND1      X922      1000                                 
 UP BND1      X923      1000                                 
 UP BND1      X924      1000                                 
 LO BND1      X925      -1000                                
 UP BND1      X925      1000                                 
 UP BND1      X926      1000                                 
 UP BND1      X927      1000                                 
 LO BND1      X928      -1000                                
 UP BND1      X928      1000                                 
 LO BND1      X929      -1000                                
 UP BND1      X929      1000                                 
 UP BND1      X930      1000                                 
 UP BND1      X931      1000                                 
 UP BND1      X932      1000                                 
 UP BND1      X933      1000                                 
 UP BND1      X934      1000                                 
 UP BND1      X935      1000                                 
 UP BND1      X936      1000                                 
 UP BND1      X937      1000                                 
 UP BND1      X938      1000                                 
 LO BND1      X939      -1000                                
 UP BND1      X939      1000                                 
 LO BND1      X940      -1000                                
 UP BND1      X940      1000                                 
 LO BND1      X941      -1000                                
 UP BND1      X941      1000                                 
 LO BND1      X942      -1000                                
 UP BND1      X942      1000                                 
 LO BND1      X943      -1000                                
 UP BND1      X943      1000                                 
 LO BND1      X944      -1000                                
 UP BND1      X944      1000                                 
 UP BND1      X945      1000                                 
 LO BND1      X946      -1000                                
 UP BND1      X946      1000                                 
 LO BND1      X947      -1000                                
 UP BND1      X947      1000                                 
 UP BND1      X948      1000                                 
 UP BND1      X949      1000                                 
 UP BND1      X950      1000                                 
 UP BND1      X951      1000                                 
 LO BND1      X952      -1000                                
 UP BND1      X952      1000                                 
 LO BND1      X953      -1000                                
 UP BND1      X953      1000                                 
 UP BND1      X954      1000                                 
 UP BND1      X955      1000                                 
 UP BND1      X956      1000                                 
 UP BND1      X957      1000                                 
 LO BND1      X958      -1000                                
 UP BND1      X958      1000                                 
 UP BND1      X959      1000                                 
 LO BND1      X960      -1000                                
 UP BND1      X960      1000                                 
 UP BND1      X961      1000                                 
 LO BND1      X962      -1000                                
 UP BND1      X962      1000                                 
 LO BND1      X963      -1000                                
 UP BND1      X963      1000                                 
 UP BND1      X964      1000                                 
 LO BND1      X965      -1000                                
 UP BND1      X965      1000                                 
 LO BND1      X966      -1000                                
 UP BND1      X966      1000                                 
 LO BND1      X967      -1000                                
 UP BND1      X967      1000                                 
 UP BND1      X968      1000                                 
 UP BND1      X969      1000                                 
 UP BND1      X970      1000                                 
 UP BND1      X971      1000                                 
 LO BND1      X972      -1000                                
 UP BND1      X972      1000                                 
 UP BND1      X973      1000                                 
 UP BND1      X974      1000                                 
 LO BND1      X975      -1000                                
 UP BND1      X975      1000                                 
 LO BND1      X976      -1000                                
 UP BND1      X976      1000                                 
 LO BND1      X977      -1000                                
 UP BND1      X977      1000                                 
 LO BND1      X978      -1000                                
 UP BND1      X978      1000                                 
 UP BND1      X979      1000                                 
 UP BND1      X980      1000                                 
 LO BND1      X981      -1000                                
 UP BND1      X981      1000                                 
 UP BND1      X982      1000                                 
 LO BND1      X983      -1000                                
 UP BND1      X983      1000                                 
 UP BND1      X984      1000                                 
 LO BND1      X985      -1000                                
 UP BND1      X985      1000                                 
 LO BND1      X986      -1000                                
 UP BND1      X986      1000                                 
 UP BND1      X987      1000                                 
 LO BND1      X988      -1000                                
 UP BND1      X988      1000                                 
 UP BND1      X989      1000                                 
 UP BND1      X990      1000                                 
 LO BND1      X991      -1000                                
 UP BND1      X991      1000                                 
 LO BND1      X992      -1000                                
 UP BND1      X992      1000                                 
 LO BND1      X993      -1000                                
 UP BND1      X993      1000                                 
 LO BND1      X994      -1000                                
 UP BND1      X994      1000                                 
 UP BND1      X995      1000                                 
 UP BND1      X996      1000                                 
 LO BND1      X997      -1000                                
 UP BND1      X997      1000                                 
 LO BND1      X998      -1000                                
 UP BND1      X998      1000                                 
 UP BND1      X999      1000                                 
 LO BND1      X1000     -1000                                
 UP BND1      X1000     1000                                 
 LO BND1      X1001     -1000                                
 UP BND1      X1001     1000                                 
 UP BND1      X1002     1000                                 
 UP BND1      X1003     1000                                 
 UP BND1      X1004     1000                                 
 UP BND1      X1005     1000                                 
 LO BND1      X1006     -1000                                
 UP BND1      X1006     1000                                 
 LO BND1      X1007     -1000                                
 UP BND1      X1007     1000                                 
 LO BND1      X1008     -1000                                
 UP BND1      X1008     1000                                 
 UP BND1      X1009     1000                                 
 LO BND1      X1010     -1000                                
 UP BND1      X1010     1000                                 
 LO BND1      X1011     -1000                                
 UP BND1      X1011     1000                                 
 LO BND1      X1012     -1000                                
 UP BND1      X1012     1000                                 
 UP BND1      X1013     1000                                 
 LO BND1      X1014     -1000                                
 UP BND1      X1014     1000                                 
 UP BND1      X1015     1000                                 
 LO BND1      X1016     -1000                                
 UP BND1      X1016     1000                                 
 LO BND1      X1017     -1000                                
 UP BND1      X1017     1000                                 
 UP BND1      X1018     1000                                 
 LO BND1      X1019     -1000                                
 UP BND1      X1019     1000                                 
 LO BND1      X1020     -1000                                
 UP BND1      X1020     1000                                 
 LO BND1      X1021     -1000                                
 UP BND1      X1021     1000                                 
 UP BND1      X1022     1000                                 
 LO BND1      X1023     -1000                                
 UP BND1      X1023     1000                                 
 LO BND1      X1024     -1000                                
 UP BND1      X1024     1000                                 
 UP BND1      X1025     1000                                 
 UP BND1      X1026     1000                                 
 UP BND1      X1027     1000                                 
 UP BND1      X1028     1000                                 
 UP BND1      X1029     1000                                 
 UP BND1      X1030     1000                                 
 UP BND1      X1031     1000                                 
 UP BND1      X1032     1000                                 
 UP BND1      X1033     1000                                 
 LO BND1      X1034     -1000                                
 UP BND1      X1034     1000                                 
 LO BND1      X1035     -1000                                
 UP BND1      X1035     1000                                 
 UP BND1      X1036     1000                                 
 UP BND1      X1037     1000                                 
 UP BND1      X1038     1000                                 
 UP BND1      X1039     1000                                 
 UP BND1      X1040     1000                                 
 UP BND1      X1041     1000                                 
 UP BND1      X1042     1000                                 
 UP BND1      X1043     1000                                 
 UP BND1      X1044     1000                                 
 LO BND1      X1045     -1000                                
 UP BND1      X1045     1000                                 
 UP BND1      X1046     1000                                 
 UP BND1      X1047     1000                                 
 LO BND1      X1048     -1000                                
 UP BND1      X1048     1000                                 
 LO BND1      X1049     -1000                                
 UP BND1      X1049     1000                                 
 LO BND1      X1050     -1000                                
 UP BND1      X1050     1000                                 
 LO BND1      X1051     -1000                                
 UP BND1      X1051     1000                                 
 UP BND1      X1052     1000                                 
 UP BND1      X1053     1000                                 
 UP BND1      X1054     1000                                 
 LO BND1      X1055     -1000                                
 UP BND1      X1055     1000                                 
 UP BND1      X1056     1000                                 
 UP BND1      X1057     1000                                 
 UP BND1      X1058     1000                                 
 UP BND1      X1059     1000                                 
 UP BND1      X1060     1000                                 
ENDATA                                                       
